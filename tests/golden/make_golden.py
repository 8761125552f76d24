"""Write golden input/output vectors from the reference morphompm package.

Run in the build container only (the reference is not present elsewhere):

    python tests/golden/make_golden.py

It imports the reference from /root/reference/python and records, for small
seeded inputs, the values that no closed form or finite-difference check pins
exactly: constitutive stress/VJP values (neo-Hookean, Hencky incl. repeated
singular values, Herschel-Bulkley away from rest), single steps and their VJPs
for every model with advection on and off, 3-step rollouts and their adjoints,
dloss_dg, and the infer_growth nudge branch.  Output: ref_vectors.npz (data
only: inputs and expected outputs).
"""
import os
import sys

import numpy as np

REF_PY = "/root/reference/python"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    sys.path.insert(0, REF_PY)
    from morphompm.config import SimConfig, Material
    from morphompm.state import ParticleState
    from morphompm import constitutive as K
    from morphompm import transfer as T
    from morphompm import integrate as I
    from morphompm import diff as D

    rng = np.random.default_rng(20261203)
    out = {}
    mat = Material()

    # ---- constitutive values --------------------------------------------
    n_c = 8
    F = np.eye(3) + 0.15 * rng.standard_normal((n_c, 3, 3))
    Fg = np.eye(3) * rng.uniform(0.8, 1.3, (n_c, 1, 1)) + 0.05 * rng.standard_normal((n_c, 3, 3))
    C = 0.5 * rng.standard_normal((n_c, 3, 3))
    tb = rng.standard_normal((n_c, 3, 3))
    out["cons_F"], out["cons_Fg"], out["cons_C"], out["cons_tb"] = F, Fg, C, tb
    models = {"nh": K.NeoHookean(mat), "hencky": K.Hencky(mat), "hb": K.HerschelBulkley()}
    for name, mdl in models.items():
        tau = np.stack([mdl.stress(F[i], Fg[i], C[i]) for i in range(n_c)])
        vj = [mdl.stress_vjp(F[i], Fg[i], C[i], tb[i]) for i in range(n_c)]
        out[f"cons_{name}_tau"] = tau
        out[f"cons_{name}_Fb"] = np.stack([v[0] for v in vj])
        out[f"cons_{name}_Fgb"] = np.stack([v[1] for v in vj])
        out[f"cons_{name}_Cb"] = np.stack([v[2] for v in vj])
    # Hencky at and near repeated singular values (the _DEGEN branch)
    deg_F = np.stack([np.diag([1.5, 1.5, 1.2]) + 1e-4 * rng.standard_normal((3, 3)),
                      np.diag([1.3, 1.3, 1.3]) + 1e-4 * rng.standard_normal((3, 3)),
                      1.3 * np.eye(3),
                      np.diag([1.1, 1.1, 1.1 + 5e-10]),
                      np.diag([1.0 + 1e-6, 1.0, 1.0 - 1e-6])])
    deg_tb = rng.standard_normal((len(deg_F), 3, 3))
    hk = models["hencky"]
    out["deg_F"], out["deg_tb"] = deg_F, deg_tb
    out["deg_tau"] = np.stack([hk.stress(f, np.eye(3)) for f in deg_F])
    dv = [hk.stress_vjp(f, np.eye(3), None, t) for f, t in zip(deg_F, deg_tb)]
    out["deg_Fb"] = np.stack([v[0] for v in dv])
    out["deg_Fgb"] = np.stack([v[1] for v in dv])

    # ---- single steps + VJPs from a moving state ---------------------------
    cfg = SimConfig()
    pts = np.concatenate([np.array([[0.40, 0.40, 0.40], [0.42, 0.40, 0.41], [0.40, 0.43, 0.40],
                                    [0.41, 0.40, 0.44], [0.44, 0.44, 0.43]]),
                          rng.uniform(0.33, 0.47, (11, 3)),
                          np.array([[0.06, 0.40, 0.72]])])     # near the grid edge
    n = len(pts)
    out["step_pts"] = pts
    for name, mdl in models.items():
        for adv in (False, True):
            st = ParticleState.rest(pts)
            st.Fg[:] = (1.25 * np.eye(3) + 0.02 * rng.standard_normal((n, 3, 3)))
            st, _ = I.rollout(st, mdl, cfg, 2, advect=True)        # v, C, F non-trivial
            key = f"step_{name}_{int(adv)}"
            for fld in ("x", "v", "C", "F", "Fg"):
                out[f"{key}_in_{fld}"] = getattr(st, fld).copy()
            ns, cache = T.step(st, mdl, cfg, advect=adv)
            for fld in ("x", "v", "C", "F", "Fg"):
                out[f"{key}_out_{fld}"] = getattr(ns, fld).copy()
            bars = (rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                    rng.standard_normal((n, 3, 3)), rng.standard_normal((n, 3, 3)),
                    rng.standard_normal((n, 3, 3)))
            for fld, b in zip(("x", "v", "C", "F", "Fg"), bars):
                out[f"{key}_bar_{fld}"] = b
            res = T.step_vjp(cache, *bars, advect=adv)
            for fld, r in zip(("x", "v", "C", "F", "Fg"), res):
                out[f"{key}_vjp_{fld}"] = r

    # ---- rollouts + dloss_dg (5-point gate scene, 3 steps) -----------------
    gpts = pts[:5]
    for name, mdl in models.items():
        for adv in (False, True):
            loss, dg = I.dloss_dg(gpts, mdl, cfg, 1.30, 3, adv)
            out[f"dloss_{name}_{int(adv)}"] = np.array([loss, dg])
            st, _ = I.rollout(I.iso_growth_state(gpts, 1.30), mdl, cfg, 3, advect=adv)
            out[f"roll_{name}_{int(adv)}_F"] = st.F
            out[f"roll_{name}_{int(adv)}_x"] = st.x

    # ---- infer_growth: nudge branch (particles off-grid -> dg == 0) --------
    far = np.array([[5.0, 5.0, 5.0], [5.1, 5.0, 5.0]])
    g, its, hist = D.infer_growth(far, models["nh"], cfg, 1.0, 2, g_init=1.2,
                                  return_history=True)
    out["nudge_pts"] = far
    out["nudge_result"] = np.array([g, its])
    out["nudge_hist"] = np.array(hist)
    # ordinary branch with Hencky + advection
    st, _ = I.rollout(I.iso_growth_state(gpts, 1.25), models["hencky"], cfg, 3, advect=True)
    tgt = I.loss_detF(st.F)
    g, its, hist = D.infer_growth(gpts, models["hencky"], cfg, tgt, 3, g_init=1.1, advect=True,
                                  return_history=True)
    out["infer_hencky_target"] = np.array([tgt])
    out["infer_hencky_result"] = np.array([g, its])
    out["infer_hencky_hist"] = np.array(hist)

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
