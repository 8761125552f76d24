"""Pins the CPU oracle (oracle/ref_loop.py, oracle/ref_vec.py) before anything is
checked against it: golden vectors written by the reference itself, the
reference's own inverse-recovery fixture, closed forms, invariants and central
finite differences.  CPU only."""
import numpy as np
import pytest

from oracle import ref_loop as RL
from oracle import ref_vec as RV

FIELDS = ("x", "v", "C", "F", "Fg")


def _models():
    mat = RL.Material()
    return {"nh": RL.NeoHookean(mat), "hencky": RL.Hencky(mat), "hb": RL.HerschelBulkley()}


def _close(a, b, rtol=1e-11, atol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return np.allclose(a, b, rtol=rtol, atol=atol * max(1.0, float(np.max(np.abs(b)))))


# ---------------------------------------------------------------- fixtures
def test_inverse_recovery_fixture(inverse_recovery):
    """outputs/figures/inverse_recovery.csv: Gauss-Newton iterates at full precision."""
    cfg = RL.SimConfig()
    m = RL.NeoHookean(cfg.material)
    st, _ = RL.rollout(RL.iso_growth_state(RL.GATE_POINTS, 1.40), m, cfg, 3)
    target = RL.loss_detF(st.F)
    _, _, hist = RL.infer_growth(RL.GATE_POINTS, m, cfg, target, 3, 1.15, return_history=True)
    assert len(hist) == len(inverse_recovery)
    np.testing.assert_allclose(hist, inverse_recovery, rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_constitutive_golden(golden, name):
    m = _models()[name]
    F, Fg, C, tb = (golden[f"cons_{k}"] for k in ("F", "Fg", "C", "tb"))
    for i in range(len(F)):
        assert _close(m.stress(F[i], Fg[i], C[i]), golden[f"cons_{name}_tau"][i])
        Fb, Fgb, Cb = m.stress_vjp(F[i], Fg[i], C[i], tb[i])
        assert _close(Fb, golden[f"cons_{name}_Fb"][i])
        assert _close(Fgb, golden[f"cons_{name}_Fgb"][i])
        assert _close(Cb, golden[f"cons_{name}_Cb"][i])
    # vectorised form, same vectors
    assert _close(RV.stress(m, F, Fg, C), golden[f"cons_{name}_tau"])
    Fb, Fgb, Cb = RV.stress_vjp(m, F, Fg, C, tb)
    assert _close(Fb, golden[f"cons_{name}_Fb"])
    assert _close(Fgb, golden[f"cons_{name}_Fgb"])
    assert _close(Cb, golden[f"cons_{name}_Cb"])


def test_hencky_degenerate_golden(golden):
    m = _models()["hencky"]
    I3 = np.eye(3)
    for i, (F, tb) in enumerate(zip(golden["deg_F"], golden["deg_tb"])):
        assert _close(m.stress(F, I3), golden["deg_tau"][i], rtol=1e-10)
        Fb, Fgb, _ = m.stress_vjp(F, I3, None, tb)
        assert _close(Fb, golden["deg_Fb"][i], rtol=1e-9, atol=1e-10)
        assert _close(Fgb, golden["deg_Fgb"][i], rtol=1e-9, atol=1e-10)
    Fb, Fgb, _ = RV.stress_vjp(m, golden["deg_F"], np.repeat(I3[None], len(golden["deg_F"]), 0),
                               None, golden["deg_tb"])
    assert _close(Fb, golden["deg_Fb"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("adv", [0, 1])
@pytest.mark.parametrize("impl", ["loop", "vec"])
def test_step_and_vjp_golden(golden, name, adv, impl):
    key = f"step_{name}_{adv}"
    cfg = RL.SimConfig()
    m = _models()[name]
    st = RL.ParticleState(*(golden[f"{key}_in_{f}"].copy() for f in FIELDS))
    bars = [golden[f"{key}_bar_{f}"] for f in FIELDS]
    mod = RL if impl == "loop" else RV
    ns, cache = mod.step(st, m, cfg, advect=bool(adv))
    for f in FIELDS:
        assert _close(getattr(ns, f), golden[f"{key}_out_{f}"]), f
    res = mod.step_vjp(cache, *bars, advect=bool(adv))
    for f, r in zip(FIELDS, res):
        assert _close(r, golden[f"{key}_vjp_{f}"], rtol=1e-10, atol=1e-11), f


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("adv", [0, 1])
def test_dloss_and_rollout_golden(golden, name, adv):
    cfg = RL.SimConfig()
    m = _models()[name]
    pts = golden["step_pts"][:5]
    for mod in (RL, RV):
        loss, dg = mod.dloss_dg(pts, m, cfg, 1.30, 3, bool(adv))
        np.testing.assert_allclose([loss, dg], golden[f"dloss_{name}_{adv}"], rtol=1e-10)
        st, _ = mod.rollout(mod.iso_growth_state(pts, 1.30), m, cfg, 3, advect=bool(adv))
        assert _close(st.F, golden[f"roll_{name}_{adv}_F"])
        assert _close(st.x, golden[f"roll_{name}_{adv}_x"])


def test_infer_growth_nudge_and_hencky_golden(golden):
    cfg = RL.SimConfig()
    mods = _models()
    g, its, hist = RL.infer_growth(golden["nudge_pts"], mods["nh"], cfg, 1.0, 2, 1.2,
                                   return_history=True)
    np.testing.assert_allclose([g, its], golden["nudge_result"], rtol=1e-15)
    np.testing.assert_allclose(hist, golden["nudge_hist"], rtol=1e-15)
    g, its, hist = RL.infer_growth(golden["step_pts"][:5], mods["hencky"], cfg,
                                   float(golden["infer_hencky_target"][0]), 3, 1.1,
                                   advect=True, return_history=True)
    np.testing.assert_allclose([g, its], golden["infer_hencky_result"], rtol=1e-12)
    np.testing.assert_allclose(hist, golden["infer_hencky_hist"], rtol=1e-12)


# ------------------------------------------------------------- closed forms
def test_closed_forms_T1_T1b_confined_swell():
    """tests/test_growth.cpp:56-92 (T1, T1b) and constitutive.py:155-179."""
    mat = RL.Material()
    mu, lam = mat.mu, mat.lam
    hk = RL.Hencky(mat)
    g = 1.5
    tau = hk.stress(np.eye(3), g * np.eye(3))
    np.testing.assert_allclose(np.diag(tau), -(2 * mu + 3 * lam) * np.log(g), rtol=1e-12)
    assert abs(tau[1, 2]) < 1e-9
    g1, g2, g3 = 1.2, 1.0, 0.8
    e = -np.log([g1, g2, g3])
    tau = hk.stress(np.eye(3), np.diag([g1, g2, g3]))
    np.testing.assert_allclose(np.diag(tau), 2 * mu * e + lam * e.sum(), rtol=1e-12, atol=1e-9)
    for m in (RL.NeoHookean(mat), hk):
        Fg = np.diag([1.3, 1.1, 0.9])
        assert np.max(np.abs(m.stress(Fg.copy(), Fg.copy()))) < 1e-8
    lz, txx = RL.confined_swell_neohookean(mat, 1.5)
    tau = RL.NeoHookean(mat).stress(np.diag([1.0, 1.0, lz]), 1.5 * np.eye(3))
    assert abs(tau[2, 2]) < 1e-6 and abs(tau[0, 0] - txx) < 1e-6 * abs(txx) + 1e-6
    assert abs(RL.HerschelBulkley().stress(np.eye(3), np.eye(3), np.zeros((3, 3)))).max() < 1e-8


def test_objectivity_T4():
    """tests/test_growth.cpp:173-188: tau(R F) = R tau(F) R^T."""
    mat = RL.Material()
    F0 = np.array([[1.2, 0.0, 0.03], [0.10, 0.9, 0.0], [0.0, 0.05, 1.1]])
    Fg = np.diag([1.3, 1.0, 0.8])
    ax = np.array([0.3, 0.7, 0.6]); ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(0.9) * K + (1 - np.cos(0.9)) * K @ K
    for m in (RL.NeoHookean(mat), RL.Hencky(mat)):
        t0 = m.stress(F0, Fg)
        np.testing.assert_allclose(m.stress(R @ F0, Fg), R @ t0 @ R.T, atol=1e-9 * abs(t0).max())


def test_bspline_identities():
    """sum w = 1, sum w dpos = 0, sum w dpos dpos^T = dx^2/4 I, sum grad w = 0."""
    rng = np.random.default_rng(1)
    dx = 0.037
    for _ in range(50):
        xp = rng.uniform(0.2, 0.8, 3)
        base, fx, w, dw = RL.bspline(xp, 1.0 / dx)
        assert np.all((fx >= 0.5) & (fx < 1.5))
        S0 = 0.0; S1 = np.zeros(3); S2 = np.zeros((3, 3)); Sg = np.zeros(3)
        for (i, j, k) in RL.STENCIL:
            wt = w[i, 0] * w[j, 1] * w[k, 2]
            d = (np.array([i, j, k]) - fx) * dx
            S0 += wt; S1 += wt * d; S2 += wt * np.outer(d, d)
            Sg += RL.grad_w(i, j, k, w, dw, 1.0 / dx)
        assert abs(S0 - 1) < 1e-14
        assert np.abs(S1).max() < 1e-15
        np.testing.assert_allclose(S2, dx * dx / 4 * np.eye(3), atol=1e-16)
        assert np.abs(Sg).max() < 1e-12


def test_p2g_conservation_and_apic_exactness():
    """Grid mass = n pmass, grid momentum = sum m v (stress impulse sums to 0),
    and an affine velocity field v = A x round-trips through P2G/G2P exactly."""
    rng = np.random.default_rng(2)
    cfg = RL.SimConfig(N=16, dx=1 / 16, dt=1e-4, damping=0.0, sp=1 / 32)
    n = 200
    x = rng.uniform(0.3, 0.7, (n, 3))
    A = 0.3 * rng.standard_normal((3, 3))
    st = RL.ParticleState.rest(x)
    st.v = x @ A.T
    st.C = np.repeat(A[None], n, 0)
    st.Fg = np.repeat(1.2 * np.eye(3)[None], n, 0)
    m = RL.NeoHookean(cfg.material)
    ns, cache = RV.step(st, RL.Hencky(cfg.material), cfg, advect=False)
    gm = cache["gm"]
    assert abs(gm.sum() - n * cfg.pmass) < 1e-12 * n * cfg.pmass
    # momentum: rebuild sum_i gmom_i = sum_i gv_i gm_i / s with s = 1
    mom = (cache["gv"] * gm[:, None]).sum(0)
    np.testing.assert_allclose(mom, cfg.pmass * st.v.sum(0), rtol=1e-10, atol=1e-14)
    # APIC: with zero stress (F = Fg) the affine field is reproduced
    st2 = RL.ParticleState(x.copy(), st.v.copy(), st.C.copy(),
                           np.repeat(np.eye(3)[None], n, 0), np.repeat(np.eye(3)[None], n, 0))
    ns2, _ = RV.step(st2, m, cfg, advect=False)
    np.testing.assert_allclose(ns2.v, st.v, atol=1e-12)
    np.testing.assert_allclose(ns2.C, st.C, atol=1e-10)


# ------------------------------------------------- extensions (FD-pinned)
def test_noncubic_reduces_to_cubic():
    cfg = RL.SimConfig()
    cfg2 = RL.SimConfig(dims=(8, 8, 8))
    m = RL.Hencky(cfg.material)
    st = RL.iso_growth_state(RL.GATE_POINTS, 1.2)
    a, _ = RV.step(st, m, cfg, advect=True)
    b, _ = RV.step(st, m, cfg2, advect=True)
    assert np.array_equal(a.F, b.F) and np.array_equal(a.x, b.x)
    # a non-cubic grid that covers the particles identically gives identical results
    cfg3 = RL.SimConfig(dims=(8, 9, 10))
    c, _ = RV.step(st, m, cfg3, advect=True)
    assert np.array_equal(a.F, c.F)


def _fd_rollout(model, cfg, st0, n_steps, advect, sched, loss_fn, seed_fn, direction):
    """central FD of loss along a perturbation of G (the growth tensor)."""
    eps = 1e-6

    def run(Gp):
        s = RL.ParticleState(st0.x.copy(), st0.v.copy(), st0.C.copy(), st0.F.copy(), Gp)
        fin, _ = RV.rollout(s, model, cfg, n_steps, advect=advect, sched=sched)
        return loss_fn(fin)
    fd = (run(st0.Fg + eps * direction) - run(st0.Fg - eps * direction)) / (2 * eps)
    fin, caches = RV.rollout(st0, model, cfg, n_steps, advect=advect, sched=sched)
    bars = seed_fn(fin)
    Gb = RV.rollout_vjp(caches, *bars, advect=advect, sched=sched)[4]
    return float(np.sum(Gb * direction)), fd


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("advect", [False, True])
def test_schedule_and_position_loss_fd(name, advect):
    rng = np.random.default_rng(3)
    cfg = RL.SimConfig(N=10, dims=(10, 9, 8), dx=0.1, dt=1e-3, sp=0.05)
    model = _models()[name]
    n = 12
    st0 = RL.ParticleState.rest(rng.uniform(0.3, 0.55, (n, 3)))
    st0.Fg = 1.15 * np.eye(3) + 0.02 * rng.standard_normal((n, 3, 3))
    target = st0.x + 0.01 * rng.standard_normal((n, 3))
    sched = np.linspace(0.25, 1.0, 4)

    def loss_fn(s):
        return float(np.sum((s.x - target) ** 2) + np.sum(np.linalg.det(s.F)))

    def seed_fn(s):
        z3 = np.zeros((n, 3)); z33 = np.zeros((n, 3, 3))
        Fb = np.linalg.det(s.F)[:, None, None] * np.swapaxes(np.linalg.inv(s.F), 1, 2)
        return (2 * (s.x - target), z3, z33, Fb, z33)

    direction = rng.standard_normal((n, 3, 3))
    an, fd = _fd_rollout(model, cfg, st0, 4, advect, sched, loss_fn, seed_fn, direction)
    assert abs(an - fd) <= 1e-6 * abs(fd) + 1e-9, (an, fd)


def test_schedule_of_ones_is_reference_semantics():
    cfg = RL.SimConfig()
    m = RL.NeoHookean(cfg.material)
    pts = RL.GATE_POINTS
    a, ca = RV.rollout(RV.iso_growth_state(pts, 1.3), m, cfg, 3, advect=True)
    b, cb = RV.rollout(RV.iso_growth_state(pts, 1.3), m, cfg, 3, advect=True, sched=np.ones(3))
    assert np.array_equal(a.F, b.F)
    n = len(pts)
    bars = (np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), a.F, np.zeros((n, 3, 3)))
    ga = RV.rollout_vjp(ca, *bars, advect=True)[4]
    gb = RV.rollout_vjp(cb, *bars, advect=True, sched=np.ones(3))[4]
    np.testing.assert_allclose(ga, gb, rtol=1e-13, atol=1e-16)


def test_hb_theta_fd():
    """(tau_y, K) cotangents through a 3-step trajectory vs central FD."""
    cfg = RL.SimConfig()
    pts = RL.GATE_POINTS
    n = len(pts)

    def run(ty, K):
        m = RL.HerschelBulkley(tau_y=ty, K=K)
        fin, caches = RV.rollout(RV.iso_growth_state(pts, 1.3), m, cfg, 3, advect=True)
        return fin, caches
    fin, caches = run(50.0, 1.0)
    z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
    Fb = np.linalg.det(fin.F)[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
    _, th = RV.rollout_vjp(caches, fin.x, z3, z33, Fb, z33, advect=True, want_theta=True)

    def L(ty, K):
        f, _ = run(ty, K)
        return 0.5 * np.sum(f.x ** 2) + np.sum(np.linalg.det(f.F))
    for i, (d_ty, d_K) in enumerate([(1e-4, 0.0), (0.0, 1e-6)]):
        fd = (L(50 + d_ty, 1 + d_K) - L(50 - d_ty, 1 - d_K)) / (2 * (d_ty + d_K))
        assert abs(th[i] - fd) <= 1e-5 * abs(fd) + 1e-12, (i, th[i], fd)


# ------------------------------------------- a planted mistake is caught
def test_planted_mistakes_fail_the_pins(golden, monkeypatch):
    """Swapped Lame constants and a dropped advection term must fail the pins
    above (run here on the CPU oracle only)."""
    mat = RL.Material()
    swapped = RL.NeoHookean(mat)
    swapped.mu, swapped.lam = mat.lam, mat.mu
    lz, txx = RL.confined_swell_neohookean(mat, 1.5)
    tau = swapped.stress(np.diag([1.0, 1.0, lz]), 1.5 * np.eye(3))
    assert abs(tau[0, 0] - txx) > 1e-3 * abs(txx)
    # drop the k4 * <C_bar, v dpos> grad-w term of the G2P adjoint
    key = "step_nh_1"
    cfg = RL.SimConfig()
    st = RL.ParticleState(*(golden[f"{key}_in_{f}"].copy() for f in FIELDS))
    _, cache = RV.step(st, _models()["nh"], cfg, advect=True)
    real = np.einsum
    def broken(spec, *ops, **kw):
        if spec == "nab,nsa,nsb->ns":
            return 0 * real(spec, *ops, **kw)
        return real(spec, *ops, **kw)
    monkeypatch.setattr(RV.np, "einsum", broken)
    xb = RV.step_vjp(cache, *[golden[f"{key}_bar_{f}"] for f in FIELDS], advect=True)[0]
    monkeypatch.undo()
    assert not _close(xb, golden[f"{key}_vjp_x"], rtol=1e-6, atol=1e-8)


def test_checkpointed_det_loss_grad_equals_stored_caches():
    """ref_vec.det_loss_grad_checkpointed (checkpoint every K steps, recompute
    per segment) gives rollout + rollout_vjp's loss, growth gradient and final
    state bit for bit, with and without a schedule, for segment lengths that
    do and do not divide the step count."""
    from morphompm_amd.scenes import batch_instances
    sc = batch_instances(count=2, seed=5, n_lo=200, n_hi=400, n_steps=7)
    for s in sc:
        m = RL.NeoHookean(RL.Material()) if s.model.kind == "nh" else RL.Hencky(RL.Material())
        cfg = RL.SimConfig(N=s.cfg.N, dx=s.cfg.dx, dt=s.cfg.dt, damping=s.cfg.damping,
                           sp=s.cfg.sp, material=RL.Material(), dims=s.cfg.dims)
        for sched in (s.sched, None):
            st0 = RL.ParticleState.rest(s.pts)
            st0.Fg = np.eye(3) + 0.1 * s.dG
            fin, caches = RV.rollout(st0, m, cfg, 7, advect=True, sched=sched)
            det = np.linalg.det(fin.F)
            n = s.n
            Fb = det[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
            z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
            Gb = RV.rollout_vjp(caches, z3, z3, z33, Fb, z33, advect=True, sched=sched)[4]
            for K in (1, 3, 7):
                L, g, f2 = RV.det_loss_grad_checkpointed(st0, m, cfg, 7, s.dG, True, sched, K)
                assert L == det.sum() and g == float(np.sum(Gb * s.dG))
                assert np.array_equal(f2.F, fin.F) and np.array_equal(f2.x, fin.x)
