"""GPU edge cases against the oracle (HIP kernels only): particles straddling
every face of the grid, particles off the grid (near and very far), the grid
mass threshold, an empty state, a single particle, the largest grid axis the
cell packing allows, and a batch holding an empty and a one-particle instance
beside ordinary ones.  FP64 within 1e-12 (forward) and 1e-10 (VJP) of the
FP64 oracle, as in the other parity tests."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import ref_loop as RL
from oracle import ref_vec as RV

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "C", "F", "Fg")
I3 = np.eye(3)
STRADDLE = [(0.03, 0.455, 0.74), (0.785, 0.012, 0.33), (-0.07, 0.81, 0.5), (0.42, 0.37, 0.842)]
OFF_GRID = [(-0.4, 0.3, 0.3), (0.4, 0.95, 0.4), (1e7, 0.4, 0.4)]


def _models():
    from morphompm_amd import constitutive as K
    from morphompm_amd.config import Material
    mat = Material()
    gpu = {"nh": K.NeoHookean(mat), "hencky": K.Hencky(mat), "hb": K.HerschelBulkley()}
    ref = {"nh": RL.NeoHookean(RL.Material()), "hencky": RL.Hencky(RL.Material()),
           "hb": RL.HerschelBulkley()}
    return gpu, ref


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _cfgs(**kw):
    from morphompm_amd.config import Material, SimConfig
    base = dict(N=8, dx=0.1, dt=1e-3, damping=0.05, sp=0.02)
    base.update({k: v for k, v in kw.items() if k in ("N", "dx", "dt", "damping", "sp", "dims")})
    dens = kw.get("density", 1000.0)
    return (SimConfig(material=Material(density=dens), **base),
            RL.SimConfig(material=RL.Material(density=dens), **base))


def _edge_state(seed=11, n_inner=40):
    rng = np.random.default_rng(seed)
    pts = np.concatenate([np.array(STRADDLE), np.array(OFF_GRID),
                          rng.uniform(-0.08, 0.86, (n_inner, 3))])
    n = len(pts)
    return (pts, 0.3 * rng.standard_normal((n, 3)), 0.5 * rng.standard_normal((n, 3, 3)),
            I3 + 0.05 * rng.standard_normal((n, 3, 3)),
            1.1 * I3 + 0.03 * rng.standard_normal((n, 3, 3)))


def _step_both(arrs, name, adv, cfgs, seed=3):
    from morphompm_amd import transfer
    from morphompm_amd.state import ParticleState
    gm, rm = _models()
    gcfg, rcfg = cfgs
    ns, cache = transfer.step(ParticleState(*(a.copy() for a in arrs)), gm[name], gcfg,
                              advect=bool(adv))
    rs, rcache = RV.step(RL.ParticleState(*(a.copy() for a in arrs)), rm[name], rcfg,
                         advect=bool(adv))
    rng = np.random.default_rng(seed)
    bars = [rng.standard_normal(np.shape(a)) for a in arrs]
    gv = transfer.step_vjp(cache, *bars, advect=bool(adv))
    rv = RV.step_vjp(rcache, *bars, advect=bool(adv))
    return ns, rs, gv, rv


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("adv", [0, 1])
def test_edge_scene_step_and_vjp_fp64(name, adv):
    ns, rs, gv, rv = _step_both(_edge_state(), name, adv, _cfgs())
    for f in FIELDS:
        assert _rel(getattr(ns, f), getattr(rs, f)) < 1e-12, f
    # off-grid particles: x and F kept, v = C = 0 (exactly)
    k = len(STRADDLE)
    off = slice(k, k + len(OFF_GRID))
    x0, _, _, F0, _ = _edge_state()
    assert np.all(ns.v[off] == 0) and np.all(ns.C[off] == 0)
    assert np.array_equal(ns.F[off], F0[off]) and np.array_equal(ns.x[off], x0[off])
    for f, a, b in zip(FIELDS, gv, rv[:5]):
        if not adv and f == "x":
            assert np.all(a == 0)
            continue
        assert _rel(a, b) < 1e-10, f


@pytest.mark.parametrize("dtype, tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_grid_mass_threshold(dtype, tol):
    """One stress-free particle of mass 1e-11: nodes with w pmass <= 1e-12 are
    dropped, v' = s v0 sum(w > 0.1) (no weight within 0.1% of the threshold)."""
    from morphompm_amd import transfer
    from morphompm_amd.state import ParticleState
    gcfg, rcfg = _cfgs(density=1e-11 / 0.02 ** 3)
    x = np.array([[0.413, 0.362, 0.447]])
    v0 = np.array([[0.3, -0.2, 0.5]])
    st = [x, v0, np.zeros((1, 3, 3)), I3[None].copy(), I3[None].copy()]
    ns, _ = transfer.step(ParticleState(*(a.astype(dtype) for a in st)), _models()[0]["nh"],
                          gcfg, advect=True)
    rs, _ = RV.step(RL.ParticleState(*st), _models()[1]["nh"], rcfg, advect=True)
    full = (1 - rcfg.damping) * v0
    assert not np.allclose(rs.v, full, rtol=1e-3)          # the threshold bites
    assert _rel(ns.v, rs.v) < tol and _rel(ns.C, rs.C) < 10 * tol


def test_empty_state_step_and_vjp():
    from morphompm_amd import transfer
    from morphompm_amd.state import ParticleState
    gcfg, _ = _cfgs()
    z3, z33 = np.zeros((0, 3)), np.zeros((0, 3, 3))
    ns, cache = transfer.step(ParticleState(z3, z3, z33, z33, z33), _models()[0]["nh"], gcfg)
    assert ns.x.shape == (0, 3) and ns.F.shape == (0, 3, 3)
    res = transfer.step_vjp(cache, z3, z3, z33, z33, z33)
    assert [r.shape for r in res] == [(0, 3), (0, 3), (0, 3, 3), (0, 3, 3), (0, 3, 3)]


@pytest.mark.parametrize("x", STRADDLE[:2] + [(0.4, 0.4, 0.4)])
def test_single_particle_fp64(x):
    rng = np.random.default_rng(2)
    arrs = (np.array([x]), 0.3 * rng.standard_normal((1, 3)), 0.5 * rng.standard_normal((1, 3, 3)),
            I3[None] + 0.05 * rng.standard_normal((1, 3, 3)), 1.1 * I3[None].copy())
    ns, rs, gv, rv = _step_both(arrs, "hencky", 1, _cfgs())
    for f in FIELDS:
        assert _rel(getattr(ns, f), getattr(rs, f)) < 1e-12, f
    for f, a, b in zip(FIELDS, gv, rv[:5]):
        assert _rel(a, b) < 1e-10, f


def test_largest_axis_grid_fp64():
    """A 1022 x 4 x 4 node grid, the largest axis the 10-bit cell packing holds,
    with particles straddling and past its far end."""
    from morphompm_amd.engine import MAX_AXIS_NODES
    dx = 0.01
    L = MAX_AXIS_NODES * dx
    rng = np.random.default_rng(4)
    xs = np.concatenate([L - dx * rng.uniform(-0.9, 3.0, 24), dx * rng.uniform(0.5, 500.0, 8)])
    pts = np.stack([xs, rng.uniform(0.01, 0.03, len(xs)), rng.uniform(0.01, 0.03, len(xs))], 1)
    n = len(pts)
    arrs = (pts, 0.1 * rng.standard_normal((n, 3)), 0.1 * rng.standard_normal((n, 3, 3)),
            I3 + 0.02 * rng.standard_normal((n, 3, 3)), 1.05 * np.repeat(I3[None], n, 0))
    cfgs = _cfgs(dx=dx, dims=(MAX_AXIS_NODES, 4, 4), sp=0.005)
    ns, rs, gv, rv = _step_both(arrs, "nh", 1, cfgs)
    assert np.max(pts[:, 0] / dx - 0.5) > MAX_AXIS_NODES - 1     # some base + 2 reach 1023
    for f in FIELDS:
        assert _rel(getattr(ns, f), getattr(rs, f)) < 1e-12, f
    for f, a, b in zip(FIELDS, gv, rv[:5]):
        assert _rel(a, b) < 1e-10, f


@pytest.mark.parametrize("order", ["random", "cells"])
def test_staged_and_direct_stencil_boxes_fp64(order):
    """G2P and the per-particle reverse kernel read the stencil nodes of a
    block of 256 consecutive slots from the block's box of nodes staged in LDS
    when it fits (640 nodes in FP64), else straight from the grid.  2048
    particles in a 7-cell cube: in random slot order every block's box is the
    whole cube (the direct path); sorted by cell, every box holds a few hundred
    nodes (the staged path).  Both match the FP64 oracle's step and VJP."""
    dx = 1.0 / 24
    rng = np.random.default_rng(5)
    n = 2048
    pts = rng.uniform(0.3, 0.6, (n, 3))
    base = np.floor(pts / dx - 0.5).astype(np.int64)
    if order == "cells":
        o = np.lexsort((base[:, 2], base[:, 1], base[:, 0]))
        pts, base = pts[o], base[o]
    vol = [int(np.prod(b.max(0) - b.min(0) + 3)) for b in base.reshape(-1, 256, 3)]
    assert (min(vol) > 640) if order == "random" else (max(vol) <= 640), vol
    arrs = (pts, 0.3 * rng.standard_normal((n, 3)), 0.2 * rng.standard_normal((n, 3, 3)),
            I3 + 0.05 * rng.standard_normal((n, 3, 3)), 1.05 * np.repeat(I3[None], n, 0))
    cfgs = _cfgs(N=24, dx=dx, dt=1e-4, sp=dx / 2)
    ns, rs, gv, rv = _step_both(arrs, "hencky", 1, cfgs)
    for f in FIELDS:
        assert _rel(getattr(ns, f), getattr(rs, f)) < 1e-12, f
    for f, a, b in zip(FIELDS, gv, rv[:5]):
        assert _rel(a, b) < 1e-10, f


# ------------------------------------------- empty and tiny batch instances
def _scenes_with_empty_and_single():
    from morphompm_amd import constitutive as K
    from morphompm_amd.scenes import batch_instances
    sc = batch_instances(count=4, seed=3, n_lo=200, n_hi=900, n_steps=10)
    for s in sc:
        s.model = K.Hencky(s.cfg.material)
        s.checkpoint_every = 4
    e = sc[1]
    sc[1] = dataclasses.replace(e, pts=e.pts[:0], G=e.G[:0], dG=e.dG[:0])
    o = sc[2]
    sc[2] = dataclasses.replace(o, pts=o.pts[5:6], G=o.G[5:6], dG=o.dG[5:6])
    return sc


def test_batch_with_empty_and_single_instances_fp64():
    from test_gpu_batched import _oracle_instance, _run_batch
    sc = _scenes_with_empty_and_single()
    assert [s.n for s in sc][1:3] == [0, 1]
    bi, L, g = _run_batch(sc, torch.float64)
    assert L[1] == 0.0 and g[1] == 0.0
    p = bi.plan
    for i, s in enumerate(sc):
        if s.n == 0:
            continue
        fin, Lo, go = _oracle_instance(s, s.extra["rate"])
        x, v, C, F = (t.cpu().numpy() for t in p.unpack_dyn(bi.final_state(), i, bi.perm))
        assert np.max(np.abs(F - fin.F)) <= 1e-10 * np.max(np.abs(fin.F)), i
        assert np.max(np.abs(x - fin.x)) <= 1e-12, i
        assert abs(L[i] - Lo) <= 1e-12 * abs(Lo), i
        assert abs(g[i] - go) <= 1e-9 * abs(go) + 1e-300, (i, g[i], go)


def test_batch_with_empty_and_single_instances_bitwise_fp32():
    """Instances beside an empty one give exactly the bits they give alone."""
    from test_gpu_batched import _run_batch
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory
    sc = _scenes_with_empty_and_single()
    bi, _, _ = _run_batch(sc, torch.float32)
    for i, s in enumerate(sc):
        if s.n == 0:
            continue
        p = BatchPlan([InstanceSpec.from_cfg(s.n, s.cfg, s.model, s.pts, center=True)])
        perm = p.cell_order([s.pts])
        tr = Trajectory(p, torch.float32, s.n_steps, True, 4, s.sched)
        dyn0 = p.pack_dyn([(s.pts, np.zeros((s.n, 3)), np.zeros((s.n, 3, 3)),
                            np.repeat(np.eye(3)[None], s.n, 0))], torch.float32, perm)
        G = p.pack_mat([np.eye(3) + s.extra["rate"] * s.dG], torch.float32, perm)
        fin = tr.forward(dyn0, G)
        a = [t.cpu() for t in p.unpack_dyn(fin, 0, perm)]
        b = [t.cpu() for t in bi.plan.unpack_dyn(bi.final_state(), i, bi.perm)]
        for u, w in zip(a, b):
            assert torch.equal(u, w), i


@pytest.mark.parametrize("adv", [0, 1])
def test_grid_record_matches_oracle_grid(adv):
    """The step's grid record (m, v per node) against the oracle's gm and gv at
    every node any particle's stencil reaches (the oracle's gm > 0), for the
    edge scene."""
    from morphompm_amd import transfer
    from morphompm_amd.state import ParticleState
    gcfg, rcfg = _cfgs()
    arrs = _edge_state()
    _, cache = transfer.step(ParticleState(*(a.copy() for a in arrs)), _models()[0]["hencky"],
                             gcfg, advect=bool(adv))
    _, rc = RV.step(RL.ParticleState(*(a.copy() for a in arrs)), _models()[1]["hencky"], rcfg,
                    advect=bool(adv))
    g = cache.grid[:int(np.prod(gcfg.shape))].cpu().numpy()
    reach = rc["gm"] > 0
    assert reach.sum() > 100
    assert _rel(g[reach, 0], rc["gm"][reach]) < 1e-12
    assert _rel(g[reach, 1:], rc["gv"][reach]) < 1e-12
