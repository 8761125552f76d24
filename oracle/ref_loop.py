"""Loop-faithful FP64 restatement of morphompm's growth step + adjoint.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker, never the
product path.

Follows, function by function (citations are reference paths):
  Material / SimConfig          python/morphompm/config.py:7-40
  ParticleState                 python/morphompm/state.py:10-31
  neo-Hookean stress / VJP      python/morphompm/constitutive.py:25-39
  Hencky stress / VJP           python/morphompm/constitutive.py:42-75
  model seam wrappers           python/morphompm/constitutive.py:79-103
  Herschel-Bulkley              python/morphompm/constitutive.py:106-147
  confined-swell closed form    python/morphompm/constitutive.py:155-179
  B-spline weights              python/morphompm/transfer.py:15-30
  step (P2G, grid, G2P)         python/morphompm/transfer.py:33-79
  step_vjp                      python/morphompm/transfer.py:82-158
  rollout / rollout_vjp / ...   python/morphompm/integrate.py:15-52
  infer_growth                  python/morphompm/diff.py:9-24

The loops keep the reference's order (particle-major, stencil lexicographic
i,j,k), so grid sums accumulate in the same order as the reference.

Pinned by: tests/golden fixtures written from the reference itself
(step/step_vjp/rollout/dloss_dg/HB/Hencky-degenerate/infer_growth values),
the reference's own inverse_recovery.csv iterates, closed forms (T1, T1b,
confined swell), objectivity, B-spline partition-of-unity identities, and
central finite differences of every VJP.
"""
from dataclasses import dataclass, field

import numpy as np

EYE3 = np.eye(3)
DEGEN = 1.0e-9          # constitutive.py:17 -- repeated-eigenvalue switch
MASS_EPS = 1.0e-12      # transfer.py:57,123 -- absolute grid-mass mask


# --------------------------------------------------------------------------
# config.py:7-40
# --------------------------------------------------------------------------
@dataclass
class Material:
    E: float = 1.0e4
    nu: float = 0.3
    density: float = 1000.0

    @property
    def mu(self):
        return self.E / (2.0 * (1.0 + self.nu))

    @property
    def lam(self):
        return self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))


@dataclass
class SimConfig:
    N: int = 8
    dx: float = 0.1
    dt: float = 1.0e-3
    damping: float = 0.05
    sp: float = 0.02
    material: Material = field(default_factory=Material)
    # Extension (not in the reference): a non-cubic node grid.  None keeps the
    # reference's cubic N^3 grid exactly.
    dims: tuple = None

    @property
    def inv_dx(self):
        return 1.0 / self.dx

    @property
    def pmass(self):
        return self.material.density * self.sp ** 3

    @property
    def vol0(self):
        return self.sp ** 3

    @property
    def shape(self):
        return tuple(self.dims) if self.dims is not None else (self.N, self.N, self.N)


# --------------------------------------------------------------------------
# state.py:10-31
# --------------------------------------------------------------------------
@dataclass
class ParticleState:
    x: np.ndarray
    v: np.ndarray
    C: np.ndarray
    F: np.ndarray
    Fg: np.ndarray

    @property
    def n(self):
        return self.x.shape[0]

    @staticmethod
    def rest(x0):
        n = x0.shape[0]
        ident = np.repeat(EYE3[None], n, axis=0)
        return ParticleState(np.array(x0, dtype=float, copy=True), np.zeros((n, 3)),
                             np.zeros((n, 3, 3)), ident.copy(), ident.copy())


# --------------------------------------------------------------------------
# constitutive.py:25-75  (per particle, one 3x3 at a time)
# --------------------------------------------------------------------------
def _sym(a):
    return 0.5 * (a + a.T)


def nh_stress(F, Fg, mu, lam):
    """constitutive.py:25-28: Fe = F Fg^-1; tau = mu (Fe Fe^T - I) + lam ln(det Fe) I."""
    Fe = F @ np.linalg.inv(Fg)
    return mu * (Fe @ Fe.T - EYE3) + lam * np.log(np.linalg.det(Fe)) * EYE3


def nh_vjp(F, Fg, mu, lam, tau_bar):
    """constitutive.py:31-39."""
    Gi = np.linalg.inv(Fg)
    Fe = F @ Gi
    tb = _sym(tau_bar)
    Fe_bar = 2.0 * mu * (tb @ Fe) + lam * np.trace(tb) * np.linalg.inv(Fe).T
    return Fe_bar @ Gi.T, -Gi.T @ F.T @ Fe_bar @ Gi.T


def hencky_stress(F, Fg, mu, lam):
    """constitutive.py:42-47: tau = U diag(2 mu ln s + lam sum ln s) U^T, Fe = U S V^T."""
    Fe = F @ np.linalg.inv(Fg)
    U, s, _ = np.linalg.svd(Fe)
    ls = np.log(s)
    h = 2.0 * mu * ls + lam * ls.sum()
    return (U * h) @ U.T


def hencky_vjp(F, Fg, mu, lam, tau_bar):
    """constitutive.py:50-75: Daleckii-Krein derivative of logm(b), b = Fe Fe^T."""
    Gi = np.linalg.inv(Fg)
    Fe = F @ Gi
    U, s, _ = np.linalg.svd(Fe)
    tb = _sym(tau_bar)
    beta = s * s
    lb = np.log(beta)
    M = mu * tb + 0.5 * lam * np.trace(tb) * EYE3
    K = U.T @ M @ U
    L = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            d = beta[i] - beta[j]
            L[i, j] = (lb[i] - lb[j]) / d if abs(d) > DEGEN else 1.0 / beta[i]
    Bbar = U @ (K * L) @ U.T
    Fe_bar = 2.0 * Bbar @ Fe
    return Fe_bar @ Gi.T, -Gi.T @ F.T @ Fe_bar @ Gi.T


class _Elastic:
    """constitutive.py:79-93: rate-independent seam; C ignored, C_bar = 0."""
    kind = None

    def __init__(self, material):
        self.mu = material.mu
        self.lam = material.lam

    def stress(self, F, Fg, C=None):
        f = nh_stress if self.kind == "nh" else hencky_stress
        return f(F, Fg, self.mu, self.lam)

    def stress_vjp(self, F, Fg, C, tau_bar):
        f = nh_vjp if self.kind == "nh" else hencky_vjp
        Fb, Fgb = f(F, Fg, self.mu, self.lam, tau_bar)
        return Fb, Fgb, np.zeros((3, 3))


class NeoHookean(_Elastic):
    kind = "nh"


class Hencky(_Elastic):
    kind = "hencky"


class HerschelBulkley:
    """constitutive.py:106-147 (Papanastasiou-regularised bioink).

    ``param_vjp`` is an extension: cotangents of (tau_y, K), which the
    reference does not return.  Pinned by central finite differences.
    """
    kind = "hb"

    def __init__(self, kappa=1.0e4, K=1.0, n=0.5, tau_y=50.0, m=100.0, eps=1e-4):
        self.kappa, self.K, self.n, self.tau_y, self.m, self.eps = kappa, K, n, tau_y, m, eps

    def _rate(self, C):
        D = _sym(C)
        return D, np.sqrt(2.0 * np.sum(D * D) + self.eps ** 2)

    def _eta(self, gd):
        return self.K * gd ** (self.n - 1) + self.tau_y * (1 - np.exp(-self.m * gd)) / gd

    def stress(self, F, Fg, C):
        Je = np.linalg.det(F) / np.linalg.det(Fg)
        D, gd = self._rate(C)
        return self.kappa * np.log(Je) * EYE3 + 2.0 * self._eta(gd) * D

    def stress_vjp(self, F, Fg, C, tau_bar):
        tb = _sym(tau_bar)
        tr = np.trace(tb)
        F_bar = self.kappa * tr * np.linalg.inv(F).T
        Fg_bar = -self.kappa * tr * np.linalg.inv(Fg).T
        D, gd = self._rate(C)
        e = np.exp(-self.m * gd)
        deta = (self.K * (self.n - 1) * gd ** (self.n - 2)
                + self.tau_y * ((self.m * e) * gd - (1 - e)) / gd ** 2)
        tD = float(np.sum(tb * D))
        C_bar = 2.0 * self._eta(gd) * tb + (4.0 * tD * deta / gd) * D
        return F_bar, Fg_bar, C_bar

    def param_vjp(self, F, Fg, C, tau_bar):
        """(d<tau_bar, tau>/d tau_y, d<tau_bar, tau>/d K)."""
        D, gd = self._rate(C)
        tD = float(np.sum(_sym(tau_bar) * D))
        return (2.0 * (1 - np.exp(-self.m * gd)) / gd * tD,
                2.0 * gd ** (self.n - 1) * tD)


def confined_swell_neohookean(mat, g=1.5):
    """constitutive.py:155-170: bisection for lambda_z with tau_zz = 0, closed-form tau_xx."""
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


# --------------------------------------------------------------------------
# transfer.py:15-30 -- quadratic B-spline, node-centred
# --------------------------------------------------------------------------
STENCIL = [(i, j, k) for i in range(3) for j in range(3) for k in range(3)]


def bspline(xp, inv_dx):
    """base = floor(x/dx - 1/2); fx = x/dx - base in [0.5, 1.5); w, dw of shape (3 nodes, 3 axes)."""
    X = xp * inv_dx
    base = np.floor(X - 0.5).astype(int)
    fx = X - base
    w = np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])
    dw = np.stack([fx - 1.5, -2.0 * (fx - 1.0), fx - 0.5])
    return base, fx, w, dw


def grad_w(i, j, k, w, dw, inv_dx):
    return inv_dx * np.array([dw[i, 0] * w[j, 1] * w[k, 2],
                              w[i, 0] * dw[j, 1] * w[k, 2],
                              w[i, 0] * w[j, 1] * dw[k, 2]])


def _in_grid(node, shape):
    return all(0 <= node[a] < shape[a] for a in range(3))


# --------------------------------------------------------------------------
# transfer.py:33-79 -- one forward step
# --------------------------------------------------------------------------
def step(state, model, cfg, advect=True):
    x, v, C, F, Fg = state.x, state.v, state.C, state.F, state.Fg
    n = state.n
    shp = cfg.shape
    inv_dx, dx, dt = cfg.inv_dx, cfg.dx, cfg.dt
    s, pmass, vol0 = 1.0 - cfg.damping, cfg.pmass, cfg.vol0
    k4 = 4.0 * inv_dx * inv_dx

    gm = np.zeros(shp)
    gmom = np.zeros(shp + (3,))
    per_particle = []
    for p in range(n):
        tau = model.stress(F[p], Fg[p], C[p])
        c = -k4 * vol0 * np.linalg.det(Fg[p])
        sk = c * tau
        base, fx, w, dw = bspline(x[p], inv_dx)
        for (i, j, k) in STENCIL:
            node = base + (i, j, k)
            if not _in_grid(node, shp):
                continue
            wt = w[i, 0] * w[j, 1] * w[k, 2]
            dpos = (np.array([i, j, k]) - fx) * dx
            gm[tuple(node)] += wt * pmass
            gmom[tuple(node)] += wt * (pmass * (v[p] + C[p] @ dpos) + (sk @ dpos) * dt)
        per_particle.append((base, fx, w, dw, tau, c, sk))

    gv = np.zeros(shp + (3,))
    for idx in np.argwhere(gm > MASS_EPS):
        t = tuple(idx)
        gv[t] = s * gmom[t] / gm[t]

    xn = x.copy()
    vn = np.zeros((n, 3))
    Cn = np.zeros((n, 3, 3))
    Fn = np.zeros((n, 3, 3))
    for p in range(n):
        base, fx, w, dw = per_particle[p][:4]
        nv = np.zeros(3)
        nC = np.zeros((3, 3))
        for (i, j, k) in STENCIL:
            node = base + (i, j, k)
            if not _in_grid(node, shp):
                continue
            wt = w[i, 0] * w[j, 1] * w[k, 2]
            dpos = (np.array([i, j, k]) - fx) * dx
            vi = gv[tuple(node)]
            nv += wt * vi
            nC += k4 * wt * np.outer(vi, dpos)
        vn[p] = nv
        Cn[p] = nC
        if advect:
            xn[p] = x[p] + dt * nv
        Fn[p] = (EYE3 + dt * nC) @ F[p]
    new = ParticleState(xn, vn, Cn, Fn, Fg.copy())
    cache = (state, cfg, model, gm, gv, per_particle, Cn)
    return new, cache


# --------------------------------------------------------------------------
# transfer.py:82-158 -- the hand-written reverse of step()
# --------------------------------------------------------------------------
def step_vjp(cache, xn_b, vn_b, Cn_b, Fn_b, Fgn_b, advect=True, theta_b=None):
    """Reverse of ``step``.  With advect=False x_b is returned as 0 (transfer.py:83-86).

    ``theta_b`` (extension): a list [tau_y_b, K_b] accumulated in place when the
    model has ``param_vjp``.
    """
    state, cfg, model, gm, gv, per_particle, Cn = cache
    x, v, C, F, Fg = state.x, state.v, state.C, state.F, state.Fg
    n = state.n
    shp = cfg.shape
    inv_dx, dx, dt = cfg.inv_dx, cfg.dx, cfg.dt
    s, pmass, vol0 = 1.0 - cfg.damping, cfg.pmass, cfg.vol0
    k4 = 4.0 * inv_dx * inv_dx

    F_b = np.zeros((n, 3, 3))
    v_b = np.zeros((n, 3))
    C_b = np.zeros((n, 3, 3))
    Fg_b = np.array(Fgn_b, dtype=float, copy=True)
    x_b = np.zeros((n, 3))
    gv_b = np.zeros(shp + (3,))

    # G2P adjoint (transfer.py:99-119): a scatter onto the grid
    for p in range(n):
        base, fx, w, dw = per_particle[p][:4]
        nv_b = vn_b[p] + (dt * xn_b[p] if advect else 0.0)
        if advect:
            x_b[p] += xn_b[p]
        Cb = Cn_b[p] + dt * (Fn_b[p] @ F[p].T)
        F_b[p] += (EYE3 + dt * Cn[p]).T @ Fn_b[p]
        for (i, j, k) in STENCIL:
            node = base + (i, j, k)
            if not _in_grid(node, shp):
                continue
            t = tuple(node)
            wt = w[i, 0] * w[j, 1] * w[k, 2]
            dpos = (np.array([i, j, k]) - fx) * dx
            vi = gv[t]
            gv_b[t] += wt * nv_b
            gv_b[t] += k4 * wt * (Cb @ dpos)
            if advect:
                gw = grad_w(i, j, k, w, dw, inv_dx)
                x_b[p] += (nv_b @ vi) * gw
                x_b[p] += k4 * float(np.sum(Cb * np.outer(vi, dpos))) * gw
                x_b[p] += -k4 * wt * (Cb.T @ vi)

    # grid adjoint (transfer.py:121-127)
    gmom_b = np.zeros(shp + (3,))
    gm_b = np.zeros(shp)
    for idx in np.argwhere(gm > MASS_EPS):
        t = tuple(idx)
        gmom_b[t] = s * gv_b[t] / gm[t]
        if advect:
            gm_b[t] = -float(gv_b[t] @ gv[t]) / gm[t]

    # P2G adjoint (transfer.py:129-156): a gather, then the constitutive VJP
    for p in range(n):
        base, fx, w, dw, tau, c, sk = per_particle[p]
        sk_b = np.zeros((3, 3))
        for (i, j, k) in STENCIL:
            node = base + (i, j, k)
            if not _in_grid(node, shp):
                continue
            t = tuple(node)
            wt = w[i, 0] * w[j, 1] * w[k, 2]
            dpos = (np.array([i, j, k]) - fx) * dx
            mb = gmom_b[t]
            v_b[p] += wt * pmass * mb
            C_b[p] += wt * pmass * np.outer(mb, dpos)
            sk_b += wt * dt * np.outer(mb, dpos)
            if advect:
                gw = grad_w(i, j, k, w, dw, inv_dx)
                Q2 = pmass * (v[p] + C[p] @ dpos) + (sk @ dpos) * dt
                x_b[p] += (mb @ Q2) * gw
                x_b[p] += gm_b[t] * pmass * gw
                x_b[p] += -(pmass * (C[p].T @ mb) + dt * (sk.T @ mb)) * wt
        tau_b = c * sk_b
        Fp_b, Fg_bar_stress, C_bar = model.stress_vjp(F[p], Fg[p], C[p], tau_b)
        F_b[p] += Fp_b
        C_b[p] += C_bar
        Fg_b[p] += Fg_bar_stress
        Fg_b[p] += float(np.sum(sk_b * tau)) * (-k4 * vol0) * np.linalg.det(Fg[p]) \
            * np.linalg.inv(Fg[p]).T
        if theta_b is not None and hasattr(model, "param_vjp"):
            ty_b, K_b = model.param_vjp(F[p], Fg[p], C[p], tau_b)
            theta_b[0] += ty_b
            theta_b[1] += K_b
    return x_b, v_b, C_b, F_b, Fg_b


# --------------------------------------------------------------------------
# integrate.py:15-52
# --------------------------------------------------------------------------
def rollout(state0, model, cfg, n_steps, advect=False):
    st = state0
    caches = []
    for _ in range(n_steps):
        st, cache = step(st, model, cfg, advect=advect)
        caches.append(cache)
    return st, caches


def rollout_vjp(caches, xf_b, vf_b, Cf_b, Ff_b, Fgf_b, advect=False):
    bars = (xf_b, vf_b, Cf_b, Ff_b, Fgf_b)
    for cache in reversed(caches):
        bars = step_vjp(cache, *bars, advect=advect)
    return bars


def iso_growth_state(pts, g):
    st = ParticleState.rest(pts)
    st.Fg[:] = g * EYE3
    return st


def iso_growth_grad(Fg_b):
    return float(sum(np.trace(Fg_b[p]) for p in range(len(Fg_b))))


def loss_detF(F):
    return float(sum(np.linalg.det(F[p]) for p in range(len(F))))


def dloss_dg(pts, model, cfg, g, n_steps, advect):
    """integrate.py:45-52: (sum det F_final, d/dg) via the composed adjoint."""
    st, caches = rollout(iso_growth_state(pts, g), model, cfg, n_steps, advect=advect)
    n = st.n
    Ff_b = np.stack([np.linalg.det(st.F[p]) * np.linalg.inv(st.F[p]).T for p in range(n)])
    z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
    Fg_b = rollout_vjp(caches, z3, z3, z33, Ff_b, z33, advect=advect)[4]
    return loss_detF(st.F), iso_growth_grad(Fg_b)


# --------------------------------------------------------------------------
# diff.py:9-24 -- damped Gauss-Newton on the scalar growth g
# --------------------------------------------------------------------------
def infer_growth(pts, model, cfg, target, n_steps, g_init, advect=False, iters=20,
                 tol=1e-9, return_history=False, grad_fn=None):
    """``grad_fn(g) -> (loss, dloss/dg)`` defaults to ``dloss_dg`` on this oracle."""
    if grad_fn is None:
        def grad_fn(gv):
            return dloss_dg(pts, model, cfg, gv, n_steps, advect)
    g = g_init
    hist = [g_init]
    it = -1
    for it in range(iters):
        loss, dg = grad_fn(g)
        resid = loss - target
        if abs(resid) < tol:
            break
        if abs(dg) < 1e-12 or not np.isfinite(dg):
            g += 1e-3 * (1.0 if resid < 0 else -1.0)
            continue
        g -= max(-0.5, min(0.5, resid / dg))
        hist.append(g)
    return (g, it + 1, hist) if return_history else (g, it + 1)


# The 5-point scene used by the reference's own gates and fixture
# (transfer.py:167-168, integrate.py:296-297, scripts/results.py:20-21).
GATE_POINTS = np.array([[0.40, 0.40, 0.40], [0.42, 0.40, 0.41], [0.40, 0.43, 0.40],
                        [0.41, 0.40, 0.44], [0.44, 0.44, 0.43]])
