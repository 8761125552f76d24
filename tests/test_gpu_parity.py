"""GPU path vs the oracle and the reference's golden vectors (HIP kernels only).

Tolerances: FP64 mode <= 1e-10 relative to the FP64 oracle (the bar for an
all-FP64 GPU mode); FP32 mode errors are stated per case and bounded as
written in each assert.
"""
import numpy as np
import pytest
import torch

from oracle import ref_loop as RL
from oracle import ref_vec as RV

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "C", "F", "Fg")


def _gpu_models():
    from morphompm_amd import constitutive as K
    from morphompm_amd.config import Material
    mat = Material()
    return {"nh": K.NeoHookean(mat), "hencky": K.Hencky(mat), "hb": K.HerschelBulkley()}


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_constitutive_seam_golden(golden, name):
    m = _gpu_models()[name]
    F, Fg, C, tb = (golden[f"cons_{k}"] for k in ("F", "Fg", "C", "tb"))
    tau = m.stress(F, Fg, C)
    assert _rel(tau, golden[f"cons_{name}_tau"]) < 1e-12
    Fb, Fgb, Cb = m.stress_vjp(F, Fg, C, tb)
    assert _rel(Fb, golden[f"cons_{name}_Fb"]) < 1e-12
    assert _rel(Fgb, golden[f"cons_{name}_Fgb"]) < 1e-12
    if name == "hb":
        assert _rel(Cb, golden[f"cons_{name}_Cb"]) < 1e-12
    # one particle at a time, as the reference calls it
    t0 = m.stress(F[0], Fg[0], C[0])
    assert t0.shape == (3, 3) and _rel(t0, golden[f"cons_{name}_tau"][0]) < 1e-12
    # FP32 kernels: error vs FP64 stated as a bound
    f32 = lambda a: a.astype(np.float32)
    tau32 = m.stress(f32(F), f32(Fg), f32(C))
    assert tau32.dtype == np.float32 and _rel(tau32, golden[f"cons_{name}_tau"]) < 2e-5
    Fb32, Fgb32, _ = m.stress_vjp(f32(F), f32(Fg), f32(C), f32(tb))
    assert _rel(Fb32, golden[f"cons_{name}_Fb"]) < 2e-5
    assert _rel(Fgb32, golden[f"cons_{name}_Fgb"]) < 2e-5


def test_hencky_degenerate_golden(golden):
    m = _gpu_models()["hencky"]
    F, tb = golden["deg_F"], golden["deg_tb"]
    I = np.repeat(np.eye(3)[None], len(F), 0)
    Fb, Fgb, _ = m.stress_vjp(F, I, None, tb)
    for i in range(len(F)):
        assert _rel(Fb[i], golden["deg_Fb"][i]) < 1e-9
    Fb32, _, _ = m.stress_vjp(F.astype(np.float32), I.astype(np.float32), None,
                              tb.astype(np.float32))
    for i in range(len(F)):
        assert _rel(Fb32[i], golden["deg_Fb"][i]) < 2e-5
        assert np.all(np.isfinite(Fb32[i]))


def test_fp32_hencky_series_on_device_matches_fp64():
    """The device's FP32 Hencky path (the matrix-log series while |Z|_F <=
    0.2, the FP32 eigen path beyond) against its FP64 eigen path, on the
    deformation families of the host-build test: exactly isotropic (repeated
    eigenvalues), then log-stretch spreads 1e-7 .. 0.25 that end past the
    series' range, with the stretch ratios of the last family straddling the
    switch.  Bounds as on the host: stress within 3e-6, F and Fg cotangents
    within 2e-6 of the largest FP64 entry."""
    m = _gpu_models()["hencky"]
    rng = np.random.default_rng(0)
    for sd in (0.0, 1e-7, 1e-4, 1e-2, 0.05, 0.1, 0.25):
        n = 400
        R = np.linalg.qr(rng.standard_normal((n, 3, 3)))[0]
        st = np.exp(sd * rng.standard_normal((n, 3)))
        F = R @ (st[:, :, None] * np.linalg.qr(rng.standard_normal((n, 3, 3)))[0])
        if sd == 0.0:
            F = np.repeat(np.eye(3)[None], n, 0) * 1.03
        Fg = np.repeat(np.eye(3)[None], n, 0) * rng.uniform(0.9, 1.1, (n, 1, 1))
        tb = rng.standard_normal((n, 3, 3))
        tb = tb + tb.transpose(0, 2, 1)
        f32 = lambda a: a.astype(np.float32)
        tau64 = m.stress(F, Fg, None)
        tau32 = m.stress(f32(F), f32(Fg), None)
        assert tau32.dtype == np.float32 and _rel(tau32, tau64) < 3e-6, sd
        Fb64, Fgb64, _ = m.stress_vjp(F, Fg, None, tb)
        Fb32, Fgb32, _ = m.stress_vjp(f32(F), f32(Fg), None, f32(tb))
        assert _rel(Fb32, Fb64) < 2e-6 and _rel(Fgb32, Fgb64) < 2e-6, sd


def test_hb_param_vjp_matches_oracle():
    m = _gpu_models()["hb"]
    rng = np.random.default_rng(5)
    F = np.eye(3) + 0.1 * rng.standard_normal((6, 3, 3))
    Fg = np.eye(3) * 1.1 + 0.02 * rng.standard_normal((6, 3, 3))
    C = 0.5 * rng.standard_normal((6, 3, 3))
    tb = rng.standard_normal((6, 3, 3))
    th = m.param_vjp(F, Fg, C, tb)
    ref = RV.param_vjp(RL.HerschelBulkley(), C, tb)
    assert _rel(th, ref) < 1e-12


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("adv", [0, 1])
def test_step_and_vjp_golden_fp64(golden, name, adv):
    from morphompm_amd import transfer
    from morphompm_amd.config import SimConfig
    from morphompm_amd.state import ParticleState
    key = f"step_{name}_{adv}"
    st = ParticleState(*(golden[f"{key}_in_{f}"].copy() for f in FIELDS))
    ns, cache = transfer.step(st, _gpu_models()[name], SimConfig(), advect=bool(adv))
    for f in FIELDS:
        assert _rel(getattr(ns, f), golden[f"{key}_out_{f}"]) < 1e-12, f
    res = transfer.step_vjp(cache, *(golden[f"{key}_bar_{f}"] for f in FIELDS), advect=bool(adv))
    for f, r in zip(FIELDS, res):
        ref = golden[f"{key}_vjp_{f}"]
        if not adv and f == "x":
            assert np.all(r == 0)
            continue
        assert _rel(r, ref) < 1e-10, f


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_step_fp32_error_bound(golden, name):
    from morphompm_amd import transfer
    from morphompm_amd.config import SimConfig
    from morphompm_amd.state import ParticleState
    key = f"step_{name}_1"
    st = ParticleState(*(golden[f"{key}_in_{f}"].astype(np.float32) for f in FIELDS))
    ns, cache = transfer.step(st, _gpu_models()[name], SimConfig(), advect=True)
    assert ns.F.dtype == np.float32
    assert _rel(ns.F, golden[f"{key}_out_F"]) < 1e-5
    assert _rel(ns.x, golden[f"{key}_out_x"]) < 1e-6
    res = transfer.step_vjp(cache, *(golden[f"{key}_bar_{f}"].astype(np.float32) for f in FIELDS),
                            advect=True)
    for f, r in zip(FIELDS, res):
        assert _rel(r, golden[f"{key}_vjp_{f}"]) < 1e-3, f


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("adv", [0, 1])
def test_dloss_dg_and_rollout_golden(golden, name, adv):
    from morphompm_amd import integrate as G
    from morphompm_amd.config import SimConfig
    pts = golden["step_pts"][:5]
    m = _gpu_models()[name]
    loss, dg = G.dloss_dg(pts, m, SimConfig(), 1.30, 3, bool(adv))
    ref = golden[f"dloss_{name}_{adv}"]
    assert abs(loss - ref[0]) <= 1e-12 * abs(ref[0])
    assert abs(dg - ref[1]) <= 1e-10 * abs(ref[1])
    st, caches = G.rollout(G.iso_growth_state(pts, 1.30), m, SimConfig(), 3, advect=bool(adv))
    assert _rel(st.F, golden[f"roll_{name}_{adv}_F"]) < 1e-12
    assert _rel(st.x, golden[f"roll_{name}_{adv}_x"]) < 1e-12


def test_inverse_recovery_fixture_on_gpu(inverse_recovery):
    """outputs/figures/inverse_recovery.csv reproduced by the GPU FP64 path."""
    from morphompm_amd import diff, integrate as G
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    m = _gpu_models()["nh"]
    st, _ = G.rollout(G.iso_growth_state(RL.GATE_POINTS, 1.40), m, cfg, 3)
    target = G.loss_detF(st.F)
    _, _, hist = diff.infer_growth(RL.GATE_POINTS, m, cfg, target, 3, 1.15, return_history=True)
    assert len(hist) == len(inverse_recovery)
    np.testing.assert_allclose(hist, inverse_recovery, rtol=1e-12, atol=0)


def test_infer_growth_nudge_golden(golden):
    from morphompm_amd import diff
    from morphompm_amd.config import SimConfig
    g, its, hist = diff.infer_growth(golden["nudge_pts"], _gpu_models()["nh"], SimConfig(), 1.0, 2,
                                     1.2, return_history=True)
    np.testing.assert_allclose([g, its], golden["nudge_result"], rtol=1e-15)
    np.testing.assert_allclose(hist, golden["nudge_hist"], rtol=1e-15)


def test_rollout_checkpointed_equals_full_and_vjp():
    """Checkpoint every 1, 2 or 3 of 7 steps: identical forward and adjoint to
    one segment, all against the oracle's stored-cache adjoint.  The segment
    lengths (7; 1; 2 2 2 1; 3 3 1) give reverse chains of odd and even length,
    whose calls alternate between the workspace's two payload areas."""
    from morphompm_amd import integrate as G
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    pts = RL.GATE_POINTS
    m = _gpu_models()["hencky"]
    n = len(pts)
    rng = np.random.default_rng(11)
    bars = (rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
            rng.standard_normal((n, 3, 3)), rng.standard_normal((n, 3, 3)),
            rng.standard_normal((n, 3, 3)))
    outs = []
    for K in (None, 1, 2, 3):
        st, h = G.rollout(G.iso_growth_state(pts, 1.2), m, cfg, 7, advect=True,
                          checkpoint_every=K)
        outs.append((st, G.rollout_vjp(h, *bars, advect=True)))
    for o in outs[1:]:
        for f in ("x", "F", "C"):
            assert np.array_equal(getattr(outs[0][0], f), getattr(o[0], f))
        for a, b in zip(outs[0][1], o[1]):
            assert np.array_equal(a, b)
    ost, oc = RV.rollout(RV.iso_growth_state(pts, 1.2), RL.Hencky(cfg.material), cfg, 7,
                         advect=True)
    ob = RV.rollout_vjp(oc, *bars, advect=True)
    assert _rel(outs[0][0].F, ost.F) < 1e-12
    for a, b in zip(outs[0][1], ob):
        assert _rel(a, b) < 1e-10


def test_rollout_with_schedule_matches_oracle():
    """A growth ramp ending below 1: the returned Fg is the last step's
    I + a (G - I), and rollout_vjp with a nonzero Fg cotangent returns the
    cotangent of the input G -- both as the oracle's rollout with the same
    schedule (FP64, <= 1e-10)."""
    from morphompm_amd import integrate as G
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    pts = RL.GATE_POINTS
    n = len(pts)
    sched = np.linspace(0.2, 0.8, 6)
    rng = np.random.default_rng(5)
    st0 = G.iso_growth_state(pts, 1.3)
    st0.Fg = st0.Fg + 0.05 * rng.standard_normal((n, 3, 3))
    bars = (rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
            rng.standard_normal((n, 3, 3)), rng.standard_normal((n, 3, 3)),
            rng.standard_normal((n, 3, 3)))
    m = _gpu_models()["nh"]
    st, h = G.rollout(st0, m, cfg, 6, advect=True, checkpoint_every=4, sched=sched)
    gb = G.rollout_vjp(h, *bars, advect=True)
    o0 = RL.ParticleState.rest(pts)
    o0.Fg = st0.Fg.copy()
    ost, oc = RV.rollout(o0, RL.NeoHookean(cfg.material), cfg, 6, advect=True, sched=sched)
    ob = RV.rollout_vjp(oc, *bars, advect=True, sched=sched)
    assert _rel(st.Fg, ost.Fg) < 1e-14
    assert _rel(st.F, ost.F) < 1e-12
    for a, b in zip(gb, ob):
        assert _rel(a, b) < 1e-10


def test_determinism_bitwise():
    """Two identical GPU runs are bit-identical (the fixed-order reductions)."""
    from morphompm_amd import integrate as G
    from morphompm_amd.config import SimConfig
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=6, ppc=8, seed=3)
    m = _gpu_models()["nh"]
    runs = []
    for _ in range(2):
        st, h = G.rollout(G.iso_growth_state(sc.pts.astype(np.float32), 1.08), m, sc.cfg, 20,
                          advect=True)
        n = st.n
        bars = (st.x, np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32), st.F,
                np.zeros((n, 3, 3), np.float32))
        runs.append((st.x.copy(), st.F.copy(), G.rollout_vjp(h, *bars, advect=True)))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    for a, b in zip(runs[0][2], runs[1][2]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_chained_steps_bitwise_equal_single_steps(name, dtype):
    """A trajectory runs its steps as chains (each G2P writes the next step's
    payload and sort keys); the same steps through transfer.step one call at a
    time (each self-contained) give bit-identical states and step records."""
    from morphompm_amd import integrate as G, transfer as T
    from morphompm_amd.scenes import free_swell_cube
    from morphompm_amd.state import ParticleState
    sc = free_swell_cube(cells=6, ppc=8, seed=5, model="hencky" if name == "hencky" else "nh")
    m = _gpu_models()[name]
    st0 = ParticleState.rest(torch.tensor(sc.pts, dtype=dtype, device="cuda"))
    st0.Fg = (1.06 * torch.eye(3, dtype=dtype, device="cuda")).expand(sc.n, 3, 3).clone()
    fin, h = G.rollout(st0, m, sc.cfg, 9, advect=True, checkpoint_every=4)
    st = st0
    caches = []
    for _ in range(9):
        st, cache = T.step(st, m, sc.cfg, advect=True)
        caches.append(cache)
    for f in ("x", "v", "C", "F"):
        assert torch.equal(getattr(fin, f), getattr(st, f)), f
    # the last step's records (grid, sort, payload of the kept segment), read
    # before the reverse pass recomputes earlier segments into the same store
    traj = h.traj
    k = traj._seg(traj.nseg - 1)[1] - 1
    assert torch.equal(traj.seg_grid[k], caches[-1].grid)
    pl = h.plan
    ns, nb, U = pl.n_slots, pl.n_bins, pl.n_tiles
    TB = int(np.prod(pl.tile))
    for a, b in ((traj.seg_sort[k], caches[-1].sort),):
        na = int(b[-1])
        assert int(a[-1]) == na
        assert torch.equal(a[:2 * ns], b[:2 * ns])                               # rank, sorted
        assert torch.equal(a[2 * ns:2 * ns + na * TB + 1], b[2 * ns:2 * ns + na * TB + 1])
        u0 = 2 * ns + nb + 1
        assert torch.equal(a[u0:u0 + U], b[u0:u0 + U])                           # unit_rank
        assert torch.equal(a[u0 + U:u0 + U + na], b[u0 + U:u0 + U + na])         # active_unit
    live = h.plan.live
    assert torch.equal(traj.seg_payload[k][live], caches[-1].payload[live])
    # reverse: chained (each VJP kernel prepares the step before) vs one
    # self-contained step_vjp call per step
    g = torch.Generator(device="cuda").manual_seed(7)
    n = sc.n
    bars = [torch.randn(n, *sh, dtype=dtype, device="cuda", generator=g)
            for sh in ((3,), (3,), (3, 3), (3, 3), (3, 3))]
    chained = G.rollout_vjp(h, *bars, advect=True)
    single = bars
    for c in reversed(caches):
        single = T.step_vjp(c, *single, advect=True)
    for a, b in zip(chained, single):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_kept_segments_equal_recompute(dtype):
    """keep_segments: the forward pass keeps the step records of its last M
    segments and the reverse pass recomputes only the others.  Final state,
    every cotangent and the HB parameter cotangents are bit-identical for
    M = 1 (recompute all but the last), 2 and all segments (no recompute),
    and stay so over repeated calls on the same trajectory."""
    from morphompm_amd import constitutive as K
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory, ROW_F
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=5, ppc=8, seed=6, n_steps=11)
    mat = sc.cfg.material
    m = K.HerschelBulkley(kappa=mat.lam + 2 * mat.mu / 3, K=10.0, n=0.6)
    p = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, m, sc.pts)])
    dyn0 = p.pack_dyn([(sc.pts, np.zeros((sc.n, 3)), np.zeros((sc.n, 3, 3)),
                        np.repeat(np.eye(3)[None], sc.n, 0))], dtype)
    G = p.pack_mat([sc.G], dtype)
    g = torch.Generator(device="cuda").manual_seed(3)
    bar = torch.randn(dyn0.shape, dtype=dtype, device="cuda", generator=g) * p.live
    res = {}
    for M in (1, 2, 3):                                  # 3 segments of 4, 4, 3 steps
        tr = Trajectory(p, dtype, sc.n_steps, True, 4, sc.sched, keep_segments=M)
        out = []
        for rep in range(2):
            fin = tr.forward(dyn0, G).clone()
            b0, Gb, th = tr.backward(bar, want_theta=True)
            out.append((fin, b0.clone(), Gb.clone(), th.clone(), tr.final().clone()))
        res[M] = out
    for M in (2, 3):
        for rep in range(2):
            for a, b in zip(res[1][rep], res[M][rep]):
                assert torch.equal(a, b), (M, rep)


def test_auto_keep_segments():
    """keep_segments="auto" sizes the kept stores from free device memory: a
    small scene keeps every segment, and the result equals the default's."""
    from morphompm_amd import constitutive as K
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=4, ppc=8, seed=2, n_steps=9)
    m = K.NeoHookean(sc.cfg.material)
    p = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, m, sc.pts)])
    dyn0 = p.pack_dyn([(sc.pts, np.zeros((sc.n, 3)), np.zeros((sc.n, 3, 3)),
                        np.repeat(np.eye(3)[None], sc.n, 0))], torch.float32)
    G = p.pack_mat([sc.G], torch.float32)
    a = Trajectory(p, torch.float32, sc.n_steps, True, 3, sc.sched, keep_segments="auto")
    b = Trajectory(p, torch.float32, sc.n_steps, True, 3, sc.sched)
    assert a.M == 3 and b.M == 1
    assert Trajectory.store_bytes(p, torch.float32, 3) == sum(
        t.numel() * t.element_size() for t in a._stores[0])
    fa, fb = a.forward(dyn0, G).clone(), b.forward(dyn0, G).clone()
    bar = torch.ones_like(fa) * p.live
    assert torch.equal(fa, fb)
    assert torch.equal(a.backward(bar)[1], b.backward(bar)[1])
