"""The per-step counting sort by cell (mpm_sort_by_cell) against a numpy
statement of its contract: units (P2G tiles) ranked in order with only the
occupied ones getting bins, particles stably sorted by (unit, bin in unit),
particles with no in-grid node in the dead unit's bin 0, padding slots in its
bin 1 at the tail of the instance's range, rank the inverse of sorted."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def expected_sort(plan, dyn_np):
    """Returns (sorted, start, unit_rank, n_active) the sort must produce."""
    T = np.array(plan.tile)
    TB = int(T.prod())
    unit_live = np.zeros(plan.n_tiles, dtype=bool)
    keys = np.full(plan.n_slots, -1, dtype=np.int64)      # (unit * TB + local)
    for i, s in enumerate(plan.specs):
        I = plan._insts[i]
        dims = np.array(s.shape) + 2
        sh = np.array([I.shift[a] for a in range(3)])
        nt = np.array([I.tx, I.ty, I.tz])
        dead = I.t_off + int(nt.prod())
        unit_live[dead] = True
        x = dyn_np[0:3, plan.p_off[i]:plan.p_off[i] + s.n].T
        inv_dx = dyn_np.dtype.type(1.0 / s.dx)
        X = x * inv_dx
        base = np.floor(X - dyn_np.dtype.type(0.5)).astype(np.int64) + np.array(s.origin)
        b = base + 2
        inside = np.all((b >= 0) & (b < dims), axis=1)
        bb = b + sh
        t = bb // T
        unit = I.t_off + (t[:, 0] * nt[1] + t[:, 1]) * nt[2] + t[:, 2]
        l = bb - t * T
        local = (l[:, 0] * T[1] + l[:, 1]) * T[2] + l[:, 2]
        k = np.where(inside, unit * TB + local, dead * TB)
        unit_live[unit[inside]] = True
        keys[plan.p_off[i]:plan.p_off[i] + s.n] = k
        padded = max(1, -(-s.n // 256)) * 256
        keys[plan.p_off[i] + s.n:plan.p_off[i] + padded] = dead * TB + 1
    unit_rank = np.where(unit_live, np.cumsum(unit_live) - 1, -1)
    n_active = int(unit_live.sum())
    ckeys = unit_rank[keys // TB] * TB + keys % TB
    order = np.argsort(ckeys, kind="stable")
    counts = np.bincount(ckeys, minlength=n_active * TB)
    start = np.concatenate([[0], np.cumsum(counts)])
    return order, start, unit_rank, n_active


def _plan_and_dyn(dtype, tile=(4, 8, 8), seed=0):
    from morphompm_amd import constitutive as K
    from morphompm_amd.config import SimConfig
    from morphompm_amd.engine import BatchPlan, InstanceSpec
    rng = np.random.default_rng(seed)
    specs, parts = [], []
    for n, N, center in ((300, 16, False), (1000, 24, True), (37, 12, False), (2048, 20, True)):
        cfg = SimConfig(N=N, dx=1.0 / N, dt=1e-4)
        pts = rng.uniform(0.25, 0.75, (n, 3))
        pts[: n // 10] = rng.uniform(-0.3, 1.3, (n // 10, 3))   # some leave the grid
        specs.append(InstanceSpec.from_cfg(n, cfg, K.NeoHookean(cfg.material), pts=pts[n // 10:],
                                           center=center))
        parts.append(pts)
    plan = BatchPlan(specs, tile=tile)
    z = lambda n, *s: np.zeros((n,) + s)
    dyn = plan.pack_dyn([(p, z(len(p), 3), z(len(p), 3, 3), np.tile(np.eye(3), (len(p), 1, 1)))
                         for p in parts], dtype)
    return plan, dyn


# every tile shape the library instantiates (test_built_tiles_are_listed)
TILES = [(4, 8, 8), (4, 4, 16), (8, 8, 4), (4, 4, 8), (4, 4, 4), (2, 4, 8)]


def test_built_tiles_are_listed():
    from morphompm_amd import _lib
    assert sorted(_lib.tile_shapes()) == sorted(TILES)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("tile", TILES)
def test_sort_by_cell_matches_contract(dtype, tile):
    plan, dyn = _plan_and_dyn(dtype, tile)
    sb = plan.new_sort()
    for rep in range(2):   # the second call reuses the cleared counters and flags
        plan.sort_by_cell(dyn, sb)
        torch.cuda.synchronize()
        s = sb.cpu().numpy().astype(np.int64)
        ns, nb, U = plan.n_slots, plan.n_bins, plan.n_tiles
        rank, srt, start = s[:ns], s[ns:2 * ns], s[2 * ns:2 * ns + nb + 1]
        unit_rank = s[2 * ns + nb + 1:2 * ns + nb + 1 + U]
        active = s[2 * ns + nb + 1 + U:2 * ns + nb + 1 + 2 * U]
        n_active = int(s[-1])
        order, start_e, ur_e, na_e = expected_sort(plan, dyn.cpu().numpy())
        assert n_active == na_e, rep
        np.testing.assert_array_equal(unit_rank, ur_e)
        np.testing.assert_array_equal(active[:n_active], np.nonzero(ur_e >= 0)[0])
        np.testing.assert_array_equal(start[:n_active * int(np.prod(tile)) + 1], start_e)
        np.testing.assert_array_equal(srt, order)
        np.testing.assert_array_equal(rank[srt], np.arange(ns))
        for i in range(plan.n_inst):     # each instance's sorted range is its slot range
            lo, hi = plan.p_off[i], plan.p_off[i + 1] if i + 1 < plan.n_inst else ns
            assert set(srt[lo:hi].tolist()) == set(range(lo, hi))


@pytest.mark.parametrize("tile", [(4, 4, 8), (8, 8, 4), (2, 4, 8)])
def test_host_bin_keys_match_kernel_keys(tile):
    """engine.bin_keys on a CUDA tensor (the key that orders slots when a plan
    packs particles) equals the key mpm_sort_by_cell gives each in-grid
    particle, read back from the sort record (bin of its sorted position,
    then that bin's unit and place), with the instance's unit offset removed."""
    from morphompm_amd.engine import bin_keys
    plan, dyn = _plan_and_dyn(torch.float64, tile, seed=4)
    sb = plan.new_sort()
    plan.sort_by_cell(dyn, sb)
    torch.cuda.synchronize()
    s = sb.cpu().numpy().astype(np.int64)
    ns, nb, U = plan.n_slots, plan.n_bins, plan.n_tiles
    TB = int(np.prod(tile))
    rank, start = s[:ns], s[2 * ns:2 * ns + nb + 1]
    active = s[2 * ns + nb + 1 + U:2 * ns + nb + 1 + 2 * U]
    n_active = int(s[-1])
    d = dyn.cpu().numpy()
    checked = 0
    for i, sp in enumerate(plan.specs):
        I = plan._insts[i]
        sl = slice(plan.p_off[i], plan.p_off[i] + sp.n)
        x = d[0:3, sl].T + np.array(sp.origin) * sp.dx        # absolute positions
        u = x / sp.dx - 0.5
        b = np.floor(u).astype(np.int64) + 2
        inside = np.all((b >= 0) & (b < np.array(sp.shape) + 2), axis=1)
        clear = np.all(np.abs(u - np.rint(u)) > 1e-9, axis=1)   # away from cell faces
        keep = inside & clear
        j = rank[sl][keep]
        k = np.searchsorted(start[:n_active * TB + 1], j, side="right") - 1
        kern = active[k // TB] * TB + k % TB - I.t_off * TB
        host = bin_keys(torch.tensor(x[keep], device="cuda"), sp.dx, sp.shape, tile,
                        [I.shift[a] for a in range(3)])
        assert host.is_cuda
        np.testing.assert_array_equal(host.cpu().numpy(), kern)
        checked += int(keep.sum())
    assert checked > 3000


def test_sort_moves_with_particles():
    """After the particles move, a second sort ranks the new set of units and
    clears the old ones (no stale flags or counts)."""
    plan, dyn = _plan_and_dyn(torch.float64)
    sb = plan.new_sort()
    plan.sort_by_cell(dyn, sb)
    dyn2 = dyn.clone()
    dyn2[0] += 0.11     # shift every particle in x
    plan.sort_by_cell(dyn2, sb)
    torch.cuda.synchronize()
    s = sb.cpu().numpy().astype(np.int64)
    ns = plan.n_slots
    order, start_e, ur_e, na_e = expected_sort(plan, dyn2.cpu().numpy())
    assert int(s[-1]) == na_e
    np.testing.assert_array_equal(s[ns:2 * ns], order)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_chained_step_output_rows(dtype):
    """mpm_step_fwd_chain's `out`: MPM_OUT_RECORD leaves the v rows of dyn_out
    as they were, MPM_OUT_NEXT leaves v and C, MPM_OUT_NONE every row; the
    rows written, the grid, the sort record and the next payload equal the
    full write bit for bit."""
    from morphompm_amd import _lib
    from morphompm_amd import constitutive as K
    from morphompm_amd.engine import BatchPlan, InstanceSpec, ROW_C, ROW_F, ROW_V, ROW_X
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=5, ppc=8, seed=2)
    m = K.NeoHookean(sc.cfg.material)
    res = {}
    for out in (_lib.OUT_FULL, _lib.OUT_RECORD, _lib.OUT_NEXT):
        p = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, m, sc.pts)])
        dyn = p.pack_dyn([(sc.pts, 0.01 * np.ones((sc.n, 3)), np.zeros((sc.n, 3, 3)),
                           np.repeat(np.eye(3)[None], sc.n, 0))], dtype)
        G = p.pack_mat([np.repeat(1.05 * np.eye(3)[None], sc.n, 0)], dtype)
        dyn_out = torch.full_like(dyn, 7.0)                 # sentinel in every row
        grid, sort = p.new_grid(dtype), p.new_sort()
        pay, pay_next = p.new_payload(dtype), p.new_payload(dtype)
        p.step_fwd_chain(dyn, G, None, None, dyn_out, grid, sort, pay, pay_next, True, False,
                         out)
        torch.cuda.synchronize()
        res[out] = (dyn_out, grid, sort, pay_next, p.live)
    # MPM_OUT_NONE: the step record alone (no G2P; dyn_out untouched)
    p = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, m, sc.pts)])
    dyn = p.pack_dyn([(sc.pts, 0.01 * np.ones((sc.n, 3)), np.zeros((sc.n, 3, 3)),
                       np.repeat(np.eye(3)[None], sc.n, 0))], dtype)
    G = p.pack_mat([np.repeat(1.05 * np.eye(3)[None], sc.n, 0)], dtype)
    dyn_none = torch.full_like(dyn, 7.0)
    grid, sort, pay = p.new_grid(dtype), p.new_sort(), p.new_payload(dtype)
    p.step_fwd_chain(dyn, G, None, None, dyn_none, grid, sort, pay, None, True, False,
                     _lib.OUT_NONE)
    torch.cuda.synchronize()
    assert torch.all(dyn_none == 7.0)
    assert torch.equal(grid, res[_lib.OUT_FULL][1]) and torch.equal(sort, res[_lib.OUT_FULL][2])
    full, live = res[_lib.OUT_FULL][0], res[_lib.OUT_FULL][4]
    assert not torch.any(full[:, live] == 7.0)
    for out, (d, g, srt, pn, _) in res.items():
        assert torch.equal(g, res[_lib.OUT_FULL][1]) and torch.equal(srt, res[_lib.OUT_FULL][2])
        assert torch.equal(pn[live], res[_lib.OUT_FULL][3][live])
        for r0, r1, written in ((ROW_X, ROW_X + 3, True),
                                (ROW_V, ROW_V + 3, out == _lib.OUT_FULL),
                                (ROW_C, ROW_C + 9, out != _lib.OUT_NEXT),
                                (ROW_F, ROW_F + 9, True)):
            rows = d[r0:r1][:, live]
            if written:
                assert torch.equal(rows, full[r0:r1][:, live]), (out, r0)
            else:
                assert torch.all(rows == 7.0), (out, r0)
