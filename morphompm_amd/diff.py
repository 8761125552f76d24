"""Growth inversion on the GPU (python/morphompm/diff.py:9-24).

``infer_growth`` keeps the reference's signature and logic: damped
Gauss-Newton on the scalar growth g, g <- g - clip((L - target)/L', +-0.5),
stop when |L - target| < tol, nudge g by +-1e-3 when |L'| < 1e-12 or L' is not
finite; returns (g, iterations[, history]).  L and L' come from the GPU
``dloss_dg``.

``batched.BatchedInverse.infer`` runs the same iteration for many instances at
once: one forward + adjoint launch sequence per iteration covers the whole
batch, and the update itself runs on the device (``mpm_gauss_newton``).
"""
import numpy as np
import torch

from .integrate import dloss_dg


def gauss_newton_update(g, loss, dg, target, tol):
    """One reference iteration (diff.py:14-23).  Returns (g', done, appended)."""
    resid = loss - target
    if abs(resid) < tol:
        return g, True, False
    if abs(dg) < 1e-12 or not np.isfinite(dg):
        return g + 1e-3 * (1.0 if resid < 0 else -1.0), False, False
    return g - max(-0.5, min(0.5, resid / dg)), False, True


def infer_growth(pts, model, cfg, target, n_steps, g_init, advect=False, iters=20, tol=1e-9,
                 return_history=False, *, dtype=torch.float64):
    g = g_init
    hist = [g_init]
    it = -1
    for it in range(iters):
        loss, dg = dloss_dg(pts, model, cfg, g, n_steps, advect, dtype=dtype)
        g, done, appended = gauss_newton_update(g, loss, dg, target, tol)
        if done:
            break
        if appended:
            hist.append(g)
    return (g, it + 1, hist) if return_history else (g, it + 1)
