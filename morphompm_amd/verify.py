"""Verification harness on the GPU path (python/morphompm/verify.py:11-38).

    python -m morphompm_amd.verify

Runs the reference pipeline's gates against the HIP backend, each with the
reference's scene and threshold, in FP64 mode (float64 numpy inputs run the
FP64 kernels; central differences need them):

  [2]  constitutive seam: stress-free states, VJP vs central FD for
       neo-Hookean, Hencky (incl. repeated singular values) and
       Herschel-Bulkley, plus closed forms (constitutive.py:150-246)
  [3]  one step of transfer.step / step_vjp vs FD (transfer.py:162-193)
  [1p] free swell expands toward det F = g^3 for every model (integrate.py:67-89)
  [4]  3-step trajectory gradient vs FD, advect off and on, elastic and
       Herschel-Bulkley (integrate.py:55-64, 113-128)
  [6]  inverse growth-rate recovery g = 1.40 from 1.15 (diff.py:27-44)
  [8]  state / config I/O round trip, state in device memory (io.py:116-148)
  [D]  two identical GPU runs are bit-identical (integrate.py:92-110)

``run()`` prints one line per check and returns 0 when every gate passes.
"""
import numpy as np

from . import constitutive as K
from . import diff, integrate, io, transfer
from .config import Material, SimConfig
from .state import ParticleState

GATE_POINTS = np.array([[0.40, 0.40, 0.40], [0.42, 0.40, 0.41], [0.40, 0.43, 0.40],
                        [0.41, 0.40, 0.44], [0.44, 0.44, 0.43]])


def _frob(a, b):
    return float(np.sum(np.asarray(a) * np.asarray(b)))


def _mark(ok):
    return "ok" if ok else "FAIL"


def _vjp_fd_err(model, F, Fg, rng, C=None, eps=1e-6):
    """Relative gap between <VJP(tau_bar), d> and the central difference of
    <tau_bar, stress> along a random direction d in (F, Fg[, C])."""
    tb = rng.standard_normal((3, 3))
    tb = 0.5 * (tb + tb.T)
    Fb, Fgb, Cb = model.stress_vjp(F, Fg, C, tb)
    dF, dG = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
    dC = rng.standard_normal((3, 3)) if C is not None else None

    def L(s):
        Cs = C + s * eps * dC if C is not None else None
        return _frob(tb, model.stress(F + s * eps * dF, Fg + s * eps * dG, Cs))
    fd = (L(1) - L(-1)) / (2 * eps)
    an = _frob(Fb, dF) + _frob(Fgb, dG) + (_frob(Cb, dC) if C is not None else 0.0)
    return abs(an - fd) / (abs(fd) + 1e-12)


def constitutive_gate():
    mat = Material()
    rng = np.random.default_rng(0)
    I3 = np.eye(3)
    states = [(np.diag([1.20, 1.00, 0.85]) + 0.03 * np.roll(I3, 1, axis=1), np.diag([1.3, 1.1, 0.9])),
              (np.diag([1.05, 0.95, 1.15]) + 0.02 * rng.standard_normal((3, 3)),
               np.diag([1.20, 0.85, 1.05])),
              (I3 + 0.05 * rng.standard_normal((3, 3)), np.diag([1.4, 1.0, 0.8]))]
    fails = 0
    for name, model in (("NeoHookean", K.NeoHookean(mat)), ("Hencky", K.Hencky(mat))):
        Fg0 = np.diag([1.3, 1.1, 0.9])
        z = float(np.max(np.abs(model.stress(Fg0.copy(), Fg0.copy()))))
        worst = max(_vjp_fd_err(model, F, Fg, rng) for F, Fg in states for _ in range(4))
        ok = z < 1e-8 and worst < 1e-5
        print(f"    [{name}] F = Fg -> |tau| = {z:.1e}; VJP vs FD worst {worst:.2e}  [{_mark(ok)}]")
        fails += not ok
    hk = K.Hencky(mat)
    deg = [(np.diag([1.5, 1.5, 1.2]) + 1e-4 * rng.standard_normal((3, 3)), I3),
           (np.diag([1.3, 1.3, 1.3]) + 1e-4 * rng.standard_normal((3, 3)), I3)]
    worst = max(_vjp_fd_err(hk, F, Fg, rng) for F, Fg in deg for _ in range(4))
    Fb, Fgb, _ = hk.stress_vjp(1.3 * I3, I3, None, I3)
    ok = worst < 1e-3 and bool(np.all(np.isfinite(Fb)) and np.all(np.isfinite(Fgb)))
    print(f"    [Hencky, repeated singular values] VJP vs FD worst {worst:.2e}; exact "
          f"isotropy finite  [{_mark(ok)}]")
    fails += not ok
    hb = K.HerschelBulkley()
    z = float(np.max(np.abs(hb.stress(I3, I3, np.zeros((3, 3))))))
    worst = max(_vjp_fd_err(hb, F, Fg, rng, C=0.5 * rng.standard_normal((3, 3)))
                for F, Fg in states for _ in range(4))
    ok = z < 1e-8 and worst < 1e-5
    print(f"    [HerschelBulkley] rest -> |tau| = {z:.1e}; VJP(F, Fg, C) vs FD worst {worst:.2e}"
          f"  [{_mark(ok)}]")
    fails += not ok
    # closed forms: confined swell (neo-Hookean) and Hencky at F = I, Fg = g I
    lz, txx = K.confined_swell_neohookean(mat, g=1.5)
    tau = K.NeoHookean(mat).stress(np.diag([1.0, 1.0, lz]), 1.5 * I3)
    ok = abs(tau[2, 2]) < 1e-6 and abs(tau[0, 0] - txx) < 1e-6 * abs(txx) + 1e-6
    g = 1.2
    th = K.Hencky(mat).stress(I3.copy(), g * I3)
    exp = -(2 * mat.mu + 3 * mat.lam) * np.log(g)
    ok &= bool(np.max(np.abs(th - exp * I3)) < 1e-9 * abs(exp))
    print(f"    [closed forms] confined swell tau_zz = {tau[2, 2]:.1e}, tau_xx = {tau[0, 0]:.3f} "
          f"vs {txx:.3f}; Hencky -(2mu+3lam) ln g  [{_mark(ok)}]")
    fails += not ok
    return fails


def transfer_gate():
    cfg = SimConfig()
    model = K.NeoHookean(cfg.material)
    n = len(GATE_POINTS)

    def loss(g):
        ns, _ = transfer.step(integrate.iso_growth_state(GATE_POINTS, g), model, cfg, advect=False)
        return float(np.sum(np.linalg.det(ns.F)))
    g0 = 1.30
    ns, cache = transfer.step(integrate.iso_growth_state(GATE_POINTS, g0), model, cfg,
                              advect=False)
    Fn_b = np.linalg.det(ns.F)[:, None, None] * np.swapaxes(np.linalg.inv(ns.F), 1, 2)
    z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
    Fg_b = transfer.step_vjp(cache, z3, z3, z33, Fn_b, z33, advect=False)[4]
    dg = float(np.trace(Fg_b, axis1=1, axis2=2).sum())
    eps = 1e-6
    fd = (loss(g0 + eps) - loss(g0 - eps)) / (2 * eps)
    rel = abs(dg - fd) / (abs(fd) + 1e-30)
    ok = rel < 1e-5
    print(f"    adjoint {dg:.6e}  FD {fd:.6e}  rel.err {rel:.2e}  [{_mark(ok)}]")
    return int(not ok)


def forward_physics_gate():
    cfg = SimConfig(N=12, dx=0.05, dt=1.0e-3, damping=0.05, sp=0.03)
    g, n_steps = 1.5, 150
    c = np.arange(0.30 - 0.015, 0.30 + 0.015 + 1e-9, 0.03)
    pts = np.array([(x, y, z) for x in c for y in c for z in c])
    fails = 0
    for name, model in (("NeoHookean", K.NeoHookean(cfg.material)),
                        ("Hencky", K.Hencky(cfg.material)), ("HerschelBulkley", K.HerschelBulkley())):
        st, _ = integrate.rollout(integrate.iso_growth_state(pts, g), model, cfg, n_steps,
                                  advect=True)
        d = float(np.mean(np.linalg.det(st.F)))
        ok = bool(np.isfinite(d) and 1.5 < d < 2.0 * g ** 3)
        print(f"    [{name}] mean det F = {d:.4f} (toward g^3 = {g ** 3:.3f})  [{_mark(ok)}]")
        fails += not ok
    return fails


def trajectory_gate():
    cfg = SimConfig()
    fails = 0
    for name, model, adv in (("NeoHookean", K.NeoHookean(cfg.material), False),
                             ("NeoHookean", K.NeoHookean(cfg.material), True),
                             ("HerschelBulkley", K.HerschelBulkley(), False)):
        g0, eps = 1.30, 1e-6
        _, dg = integrate.dloss_dg(GATE_POINTS, model, cfg, g0, 3, adv)
        fp, _ = integrate.dloss_dg(GATE_POINTS, model, cfg, g0 + eps, 3, adv)
        fm, _ = integrate.dloss_dg(GATE_POINTS, model, cfg, g0 - eps, 3, adv)
        fd = (fp - fm) / (2 * eps)
        rel = abs(dg - fd) / (abs(fd) + 1e-30)
        ok = rel < 1e-5
        print(f"    [{name}, advect={adv}] 3 steps: adjoint {dg:.6e}  FD {fd:.6e}  "
              f"rel.err {rel:.2e}  [{_mark(ok)}]")
        fails += not ok
    return fails


def inverse_gate():
    cfg = SimConfig()
    model = K.NeoHookean(cfg.material)
    g_true = 1.40
    st, _ = integrate.rollout(integrate.iso_growth_state(GATE_POINTS, g_true), model, cfg, 3)
    target = integrate.loss_detF(st.F)
    g_rec, iters = diff.infer_growth(GATE_POINTS, model, cfg, target, 3, g_init=1.15)
    err = abs(g_rec - g_true)
    ok = err < 1e-4
    print(f"    recovered g = {g_rec:.6f} in {iters} iterations (err {err:.2e})  [{_mark(ok)}]")
    return int(not ok)


def determinism_gate():
    cfg = SimConfig()
    model = K.NeoHookean(cfg.material)

    def run():
        st, _ = integrate.rollout(integrate.iso_growth_state(GATE_POINTS, 1.3), model, cfg, 3,
                                  advect=True)
        return np.array(st.F)
    d = float(np.max(np.abs(run() - run())))
    ok = d == 0.0
    print(f"    max |F_run1 - F_run2| = {d:.1e}  [{_mark(ok)}]")
    return int(not ok)


GATES = (("[2] constitutive seam: forward + VJP vs FD", constitutive_gate),
         ("[3] transfer: single-step assembly", transfer_gate),
         ("[1p] forward physics: free swell det F -> g^3", forward_physics_gate),
         ("[4] integrate: multi-step trajectory", trajectory_gate),
         ("[6] diff: inverse growth-rate recovery", inverse_gate),
         ("[8] io: round trip, state in device memory", lambda: io.main(device="cuda")),
         ("[D] determinism: identical GPU runs bit-identical", determinism_gate))


def run():
    fails = 0
    for title, gate in GATES:
        print(f"### {title} ###", flush=True)
        fails += gate()
    print("PIPELINE VERIFY (GPU):", "ALL GATES PASS" if fails == 0 else f"{fails} FAILURE(S)")
    return 0 if fails == 0 else 1


if __name__ == "__main__":
    raise SystemExit(run())
