"""Constitutive seam on the GPU (python/morphompm/constitutive.py:1-147).

Same interface as the reference:
    model.stress(F, Fg, C=None)          -> tau (Kirchhoff, 3x3 symmetric)
    model.stress_vjp(F, Fg, C, tau_bar)  -> (F_bar, Fg_bar, C_bar)
with ``F``, ``Fg``, ``C``, ``tau_bar`` either one 3x3 (as the reference) or a
batch (n, 3, 3), numpy or torch; results come back in the caller's kind.
float64 inputs run the FP64 kernels, float32 inputs the FP32 kernels.
Every call runs the HIP library (mpm_stress_* / mpm_stress_vjp_*); there is
no CPU path.

``HerschelBulkley.param_vjp`` (extension) returns the cotangents of
(tau_y, K), which the reference does not provide.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._convert import Like, to_dev

I3 = np.eye(3)
DEGEN = 1.0e-9   # constitutive.py:17; the FP64 kernels use the same switch


def _soa(a, n, dtype):
    return to_dev(a, dtype).reshape(n, 9).T.contiguous()


def _call(kind, prm, F, Fg, C, tau_bar=None, want_theta=False):
    _lib.require_gpu()
    like = Like(F)
    single = tuple(np.shape(F)) == (3, 3)
    n = 1 if single else int(np.shape(F)[0])
    dt = like.dtype
    sfx = _lib.suffix(dt)
    Fd, Gd = _soa(F, n, dt), _soa(Fg, n, dt)
    Cd = _soa(C, n, dt) if C is not None else None
    prm6 = (ctypes.c_double * 6)(*prm)
    L = _lib.lib()
    stream = _lib.stream_ptr()

    def back(t):
        m = t.T.reshape(n, 3, 3)
        return like.out(m[0] if single else m)

    if tau_bar is None:
        tau = torch.empty(9, n, dtype=dt, device="cuda")
        _lib.check(getattr(L, f"mpm_stress_{sfx}")(kind, prm6, n, n, _lib.ptr(Fd), _lib.ptr(Gd),
                                                    _lib.ptr(Cd), _lib.ptr(tau), stream), "stress")
        return back(tau)
    Tb = _soa(tau_bar, n, dt)
    Fb = torch.empty(9, n, dtype=dt, device="cuda")
    Gb = torch.empty_like(Fb)
    Cb = torch.empty_like(Fb)
    th = torch.empty(2, n, dtype=dt, device="cuda")
    _lib.check(getattr(L, f"mpm_stress_vjp_{sfx}")(kind, prm6, n, n, _lib.ptr(Fd), _lib.ptr(Gd),
                                                    _lib.ptr(Cd), _lib.ptr(Tb), _lib.ptr(Fb),
                                                    _lib.ptr(Gb), _lib.ptr(Cb), _lib.ptr(th),
                                                    stream), "stress_vjp")
    if want_theta:
        t = th.T
        return like.out(t[0] if single else t)
    return back(Fb), back(Gb), back(Cb)


class _ElasticModel:
    """Rate-independent seam (constitutive.py:79-93): C ignored, C_bar = 0."""
    kind = None
    _kind_id = None

    def __init__(self, material):
        self.mu = material.mu
        self.lam = material.lam

    def _prm(self):
        return (float(self.mu), float(self.lam), 0.0, 0.0, 0.0, 0.0)

    def stress(self, F, Fg, C=None):
        return _call(self._kind_id, self._prm(), F, Fg, None)

    def stress_vjp(self, F, Fg, C, tau_bar):
        return _call(self._kind_id, self._prm(), F, Fg, None, tau_bar)


class NeoHookean(_ElasticModel):
    """tau = mu (Fe Fe^T - I) + lam ln(det Fe) I, Fe = F Fg^-1 (constitutive.py:25-39)."""
    kind = "nh"
    _kind_id = _lib.NEO_HOOKEAN


class Hencky(_ElasticModel):
    """tau = U diag(2 mu ln s + lam sum ln s) U^T from Fe = U S V^T
    (constitutive.py:42-75); the device code eigen-decomposes b = Fe Fe^T."""
    kind = "hencky"
    _kind_id = _lib.HENCKY


class HerschelBulkley:
    """Papanastasiou-regularised yield-stress fluid (constitutive.py:106-147):
    tau = kappa ln(det F / det Fg) I + 2 eta(gdot) D, D = sym C,
    eta = K gdot^(n-1) + tau_y (1 - exp(-m gdot)) / gdot."""
    kind = "hb"

    def __init__(self, kappa=1.0e4, K=1.0, n=0.5, tau_y=50.0, m=100.0, eps=1e-4):
        self.kappa, self.K, self.n, self.tau_y, self.m, self.eps = kappa, K, n, tau_y, m, eps

    def _prm(self):
        return tuple(float(v) for v in (self.kappa, self.K, self.n, self.tau_y, self.m, self.eps))

    def stress(self, F, Fg, C):
        return _call(_lib.HERSCHEL_BULKLEY, self._prm(), F, Fg, C)

    def stress_vjp(self, F, Fg, C, tau_bar):
        return _call(_lib.HERSCHEL_BULKLEY, self._prm(), F, Fg, C, tau_bar)

    def param_vjp(self, F, Fg, C, tau_bar):
        """d<tau_bar, tau>/d(tau_y, K): shape (2,) for one particle, (n, 2) for a batch."""
        return _call(_lib.HERSCHEL_BULKLEY, self._prm(), F, Fg, C, tau_bar, want_theta=True)


def confined_swell_neohookean(mat, g=1.5):
    """Closed-form confined swelling of neo-Hookean (constitutive.py:155-170):
    lambda_z with tau_zz = 0 by bisection, and tau_xx = mu (1 - lambda_z^2)/g^2."""
    mu, L = mat.mu, mat.lam

    def f(lz):
        return mu * (lz ** 2 / g ** 2 - 1.0) + L * np.log(lz / g ** 3)

    lo, hi = 1e-3, 20.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0.0:
            hi = mid
        else:
            lo = mid
    lz = 0.5 * (lo + hi)
    return lz, mu * (1.0 - lz ** 2) / g ** 2
