"""Vectorised FP64 restatement of morphompm's growth step + adjoint.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker at sizes
the loop form cannot reach, and the timed host baseline of ``bench.py``.

Same arithmetic as ``oracle.ref_loop`` (which follows python/morphompm/
transfer.py:33-158, constitutive.py:25-147, integrate.py:15-52) evaluated over
all particles at once.  Grid sums use ``np.bincount`` over (particle, stencil)
pairs laid out particle-major, i.e. the same per-node accumulation order as the
reference's loops.

Extensions beyond the reference (each pinned by finite differences and by
reduction to the reference semantics in tests/test_oracle.py):
  * non-cubic grids ``cfg.dims = (Nx, Ny, Nz)``;
  * growth schedules: before step t, Fg_t = I + a_t (G - I) with G the state's
    Fg at t = 0 (a_t == 1 keeps G exactly);  G_bar = sum_t a_t Fg_bar_t;
  * Herschel-Bulkley parameter cotangents (tau_y, K).
"""
import numpy as np

from .ref_loop import (DEGEN, MASS_EPS, STENCIL, Material, SimConfig,  # noqa: F401
                       ParticleState, NeoHookean, Hencky, HerschelBulkley)

EYE3 = np.eye(3)
OFFS = np.array(STENCIL, dtype=np.int64)          # (27, 3), lexicographic i, j, k


def _sym(a):
    return 0.5 * (a + np.swapaxes(a, -1, -2))


def _T(a):
    return np.swapaxes(a, -1, -2)


def _tr(a):
    return np.trace(a, axis1=-2, axis2=-1)


# --------------------------------------------------------------------------
# constitutive, batched over particles (constitutive.py:25-147)
# --------------------------------------------------------------------------
def stress(model, F, Fg, C):
    kind = model.kind
    if kind == "hb":
        Je = np.linalg.det(F) / np.linalg.det(Fg)
        D = _sym(C)
        gd = np.sqrt(2.0 * np.sum(D * D, axis=(1, 2)) + model.eps ** 2)
        eta = model.K * gd ** (model.n - 1) + model.tau_y * (1 - np.exp(-model.m * gd)) / gd
        return model.kappa * np.log(Je)[:, None, None] * EYE3 + 2.0 * eta[:, None, None] * D
    mu, lam = model.mu, model.lam
    Fe = F @ np.linalg.inv(Fg)
    if kind == "nh":
        J = np.linalg.det(Fe)
        return mu * (Fe @ _T(Fe) - EYE3) + lam * np.log(J)[:, None, None] * EYE3
    U, s, _ = np.linalg.svd(Fe)
    ls = np.log(s)
    h = 2.0 * mu * ls + lam * ls.sum(axis=1, keepdims=True)
    return (U * h[:, None, :]) @ _T(U)


def stress_vjp(model, F, Fg, C, tau_bar):
    """-> (F_bar, Fg_bar, C_bar) per particle."""
    tb = _sym(tau_bar)
    kind = model.kind
    if kind == "hb":
        tr = _tr(tb)[:, None, None]
        F_bar = model.kappa * tr * _T(np.linalg.inv(F))
        Fg_bar = -model.kappa * tr * _T(np.linalg.inv(Fg))
        D = _sym(C)
        gd = np.sqrt(2.0 * np.sum(D * D, axis=(1, 2)) + model.eps ** 2)
        K, nn, ty, m = model.K, model.n, model.tau_y, model.m
        e = np.exp(-m * gd)
        eta = K * gd ** (nn - 1) + ty * (1 - e) / gd
        deta = K * (nn - 1) * gd ** (nn - 2) + ty * ((m * e) * gd - (1 - e)) / gd ** 2
        tD = np.sum(tb * D, axis=(1, 2))
        C_bar = 2.0 * eta[:, None, None] * tb + (4.0 * tD * deta / gd)[:, None, None] * D
        return F_bar, Fg_bar, C_bar
    mu, lam = model.mu, model.lam
    Gi = np.linalg.inv(Fg)
    Fe = F @ Gi
    if kind == "nh":
        Fe_bar = 2.0 * mu * (tb @ Fe) + lam * _tr(tb)[:, None, None] * _T(np.linalg.inv(Fe))
    else:
        U, s, _ = np.linalg.svd(Fe)
        beta = s * s
        lb = np.log(beta)
        M = mu * tb + 0.5 * lam * _tr(tb)[:, None, None] * EYE3
        Kt = _T(U) @ M @ U
        d = beta[:, :, None] - beta[:, None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            Ld = (lb[:, :, None] - lb[:, None, :]) / d
        L = np.where(np.abs(d) > DEGEN, Ld, 1.0 / beta[:, :, None] + 0.0 * d)
        Fe_bar = 2.0 * (U @ (Kt * L) @ _T(U)) @ Fe
    return Fe_bar @ _T(Gi), -_T(Gi) @ _T(F) @ Fe_bar @ _T(Gi), np.zeros_like(F)


def param_vjp(model, C, tau_bar):
    """Herschel-Bulkley (tau_y, K) cotangent per particle, shape (n, 2)."""
    D = _sym(C)
    gd = np.sqrt(2.0 * np.sum(D * D, axis=(1, 2)) + model.eps ** 2)
    tD = np.sum(_sym(tau_bar) * D, axis=(1, 2))
    return np.stack([2.0 * (1 - np.exp(-model.m * gd)) / gd * tD,
                     2.0 * gd ** (model.n - 1) * tD], axis=1)


# --------------------------------------------------------------------------
# stencil geometry (transfer.py:18-30,47-51)
# --------------------------------------------------------------------------
class _Stencil:
    __slots__ = ("base", "fx", "wt", "gw", "dpos", "flat", "inb")

    def __init__(self, x, cfg, need_grad):
        inv_dx, dx = cfg.inv_dx, cfg.dx
        Nx, Ny, Nz = cfg.shape
        X = x * inv_dx
        base = np.floor(X - 0.5).astype(np.int64)
        fx = X - base
        w = np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], 1)
        wi, wj, wk = w[:, OFFS[:, 0], 0], w[:, OFFS[:, 1], 1], w[:, OFFS[:, 2], 2]
        self.wt = wi * wj * wk                                       # (n, 27)
        if need_grad:
            dw = np.stack([fx - 1.5, -2.0 * (fx - 1.0), fx - 0.5], 1)
            di, dj, dk = dw[:, OFFS[:, 0], 0], dw[:, OFFS[:, 1], 1], dw[:, OFFS[:, 2], 2]
            self.gw = np.stack([di * wj * wk, wi * dj * wk, wi * wj * dk], 2) * inv_dx
        else:
            self.gw = None
        self.dpos = (OFFS[None, :, :] - fx[:, None, :]) * dx          # (n, 27, 3)
        nodes = base[:, None, :] + OFFS[None]
        self.inb = ((nodes >= 0) & (nodes < np.array([Nx, Ny, Nz]))).all(axis=2)
        self.flat = (nodes[..., 0] * Ny + nodes[..., 1]) * Nz + nodes[..., 2]
        self.base, self.fx = base, fx

    def scatter(self, vals, G):
        """sum over in-grid (p, s) pairs into G nodes; vals (n, 27[, k])."""
        idx = self.flat[self.inb]
        if vals.ndim == 2:
            return np.bincount(idx, vals[self.inb], minlength=G)
        v = vals[self.inb]
        return np.stack([np.bincount(idx, v[:, a], minlength=G) for a in range(v.shape[1])], 1)

    def gather(self, grid):
        """grid (G[, k]) -> (n, 27[, k]) with zeros for off-grid nodes."""
        safe = np.where(self.inb, self.flat, 0)
        g = grid[safe]
        if g.ndim == 3:
            return g * self.inb[..., None]
        return g * self.inb


# --------------------------------------------------------------------------
# transfer.step, vectorised (transfer.py:33-79)
# --------------------------------------------------------------------------
def step(state, model, cfg, advect=True):
    x, v, C, F, Fg = state.x, state.v, state.C, state.F, state.Fg
    G = int(np.prod(cfg.shape))
    dt, s, pmass, vol0 = cfg.dt, 1.0 - cfg.damping, cfg.pmass, cfg.vol0
    k4 = 4.0 * cfg.inv_dx * cfg.inv_dx
    st = _Stencil(x, cfg, need_grad=advect)

    tau = stress(model, F, Fg, C)
    c = -k4 * vol0 * np.linalg.det(Fg)
    sk = c[:, None, None] * tau
    Cd = np.einsum("nab,nsb->nsa", C, st.dpos)
    Sd = np.einsum("nab,nsb->nsa", sk, st.dpos)
    mom = st.wt[..., None] * (pmass * (v[:, None, :] + Cd) + Sd * dt)
    gm = st.scatter(st.wt * pmass, G)
    gmom = st.scatter(mom, G)

    act = gm > MASS_EPS
    gv = np.zeros((G, 3))
    gv[act] = s * gmom[act] / gm[act][:, None]

    vi = st.gather(gv)                                                # (n, 27, 3)
    nv = np.sum(st.wt[..., None] * vi, axis=1)
    nC = np.einsum("ns,nsa,nsb->nab", k4 * st.wt, vi, st.dpos)
    xn = x + dt * nv if advect else x.copy()
    Fn = (EYE3 + dt * nC) @ F
    new = ParticleState(xn, nv, nC, Fn, Fg.copy())
    cache = dict(state=state, cfg=cfg, model=model, gm=gm, gv=gv, tau=tau, c=c, sk=sk,
                 Cn=nC, st=st, advect=advect)
    return new, cache


# --------------------------------------------------------------------------
# transfer.step_vjp, vectorised (transfer.py:82-158)
# --------------------------------------------------------------------------
def step_vjp(cache, xn_b, vn_b, Cn_b, Fn_b, Fgn_b, advect=True, want_theta=False):
    state, cfg, model = cache["state"], cache["cfg"], cache["model"]
    x, v, C, F, Fg = state.x, state.v, state.C, state.F, state.Fg
    gm, gv, tau, c, sk, Cn, st = (cache[k] for k in ("gm", "gv", "tau", "c", "sk", "Cn", "st"))
    if advect and st.gw is None:
        raise ValueError("step_vjp(advect=True) needs a step(advect=True) cache")
    G = gm.shape[0]
    dt, s, pmass, vol0 = cfg.dt, 1.0 - cfg.damping, cfg.pmass, cfg.vol0
    k4 = 4.0 * cfg.inv_dx * cfg.inv_dx
    wt, dpos = st.wt, st.dpos

    # G2P adjoint: scatter onto the grid
    nv_b = vn_b + dt * xn_b if advect else np.array(vn_b, dtype=float, copy=True)
    Cb = Cn_b + dt * (Fn_b @ _T(F))
    F_b = _T(EYE3 + dt * Cn) @ Fn_b
    Cbd = np.einsum("nab,nsb->nsa", Cb, dpos)
    gv_b = st.scatter(wt[..., None] * nv_b[:, None, :] + k4 * wt[..., None] * Cbd, G)
    vi = st.gather(gv)
    x_b = np.zeros_like(x)
    if advect:
        x_b = np.array(xn_b, dtype=float, copy=True)
        a1 = np.einsum("na,nsa->ns", nv_b, vi)
        a2 = k4 * np.einsum("nab,nsa,nsb->ns", Cb, vi, dpos)
        x_b = x_b + np.sum((a1 + a2)[..., None] * st.gw, axis=1)
        x_b = x_b - k4 * np.einsum("ns,nba,nsb->na", wt, Cb, vi)

    # grid adjoint
    act = gm > MASS_EPS
    gmom_b = np.zeros((G, 3))
    gmom_b[act] = s * gv_b[act] / gm[act][:, None]
    gm_b = np.zeros(G)
    if advect:
        gm_b[act] = -np.sum(gv_b[act] * gv[act], axis=1) / gm[act]

    # P2G adjoint: gather + constitutive VJP
    mb = st.gather(gmom_b)
    v_b = pmass * np.sum(wt[..., None] * mb, axis=1)
    MD = np.einsum("ns,nsa,nsb->nab", wt, mb, dpos)
    C_b = pmass * MD
    sk_b = dt * MD
    if advect:
        Q2 = pmass * (v[:, None, :] + np.einsum("nab,nsb->nsa", C, dpos)) \
            + np.einsum("nab,nsb->nsa", sk, dpos) * dt
        gmb = st.gather(gm_b)
        coef = np.einsum("nsa,nsa->ns", mb, Q2) + gmb * pmass
        x_b = x_b + np.sum(coef[..., None] * st.gw, axis=1)
        x_b = x_b - np.einsum("ns,nsa->na", wt,
                              pmass * np.einsum("nba,nsb->nsa", C, mb)
                              + dt * np.einsum("nba,nsb->nsa", sk, mb))
    tau_b = c[:, None, None] * sk_b
    Fp_b, Fg_bar, C_bar = stress_vjp(model, F, Fg, C, tau_b)
    F_b = F_b + Fp_b
    C_b = C_b + C_bar
    vol_term = np.sum(sk_b * tau, axis=(1, 2)) * (-k4 * vol0) * np.linalg.det(Fg)
    Fg_b = Fgn_b + Fg_bar + vol_term[:, None, None] * _T(np.linalg.inv(Fg))
    out = (x_b, v_b, C_b, F_b, Fg_b)
    if want_theta:
        th = param_vjp(model, C, tau_b) if model.kind == "hb" else np.zeros((len(x), 2))
        return out, th
    return out


# --------------------------------------------------------------------------
# trajectories (integrate.py:15-52) + growth schedules
# --------------------------------------------------------------------------
def grown(G, a):
    """Fg = I + a (G - I); a == 1 returns G itself (no rounding)."""
    if a == 1.0:
        return G.copy()
    return EYE3 + a * (G - EYE3)


def rollout(state0, model, cfg, n_steps, advect=False, sched=None, keep=True):
    """``sched``: None (reference semantics: Fg passes through) or per-step a_t."""
    st = state0
    G = state0.Fg
    caches = []
    for t in range(n_steps):
        if sched is not None:
            st = ParticleState(st.x, st.v, st.C, st.F, grown(G, sched[t]))
        st, cache = step(st, model, cfg, advect=advect)
        if keep:
            caches.append(cache)
    return st, caches


def rollout_vjp(caches, xf_b, vf_b, Cf_b, Ff_b, Fgf_b, advect=False, sched=None,
                want_theta=False):
    """Reverse of ``rollout``.  With a schedule the returned Fg cotangent is G_bar."""
    bars = (xf_b, vf_b, Cf_b, Ff_b, Fgf_b)
    th = np.zeros(2)
    if sched is None:
        for cache in reversed(caches):
            r = step_vjp(cache, *bars, advect=advect, want_theta=want_theta)
            if want_theta:
                bars, tp = r
                th += tp.sum(axis=0)
            else:
                bars = r
    else:
        Gb = sched[len(caches) - 1] * np.asarray(Fgf_b, dtype=float)
        zero = np.zeros_like(Gb)
        for t in range(len(caches) - 1, -1, -1):
            r = step_vjp(caches[t], *bars[:4], zero, advect=advect, want_theta=want_theta)
            if want_theta:
                r, tp = r
                th += tp.sum(axis=0)
            bars = r
            Gb = Gb + sched[t] * bars[4]
        bars = bars[:4] + (Gb,)
    return (bars, th) if want_theta else bars


def iso_growth_state(pts, g):
    st = ParticleState.rest(pts)
    st.Fg[:] = g * EYE3
    return st


def loss_detF(F):
    return float(np.sum(np.linalg.det(F)))


def dloss_dg(pts, model, cfg, g, n_steps, advect):
    st, caches = rollout(iso_growth_state(pts, g), model, cfg, n_steps, advect=advect)
    n = st.n
    Ff_b = np.linalg.det(st.F)[:, None, None] * _T(np.linalg.inv(st.F))
    z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
    Fg_b = rollout_vjp(caches, z3, z3, z33, Ff_b, z33, advect=advect)[4]
    return loss_detF(st.F), float(np.sum(_tr(Fg_b)))


def det_loss_grad_checkpointed(state0, model, cfg, n_steps, dG, advect=True, sched=None, K=20):
    """L = sum_p det F_final and dL/dtheta = sum_p <dG_p, G_bar_p> for a growth
    tensor G = state0.Fg moving along dG, from ``rollout`` and ``rollout_vjp``
    restated with a checkpoint every K steps: the forward pass keeps only the
    segment-start states, the reverse pass recomputes one segment's caches at
    a time.  Same steps and the same schedule rule (G_bar = sum_t a_t Fg_bar_t)
    as rollout / rollout_vjp, so the results are theirs bit for bit, at memory
    for K caches instead of n_steps.  Returns (L, dL/dtheta, final state)."""
    G = state0.Fg

    def at(st, t):
        return ParticleState(st.x, st.v, st.C, st.F, grown(G, sched[t])) if sched is not None \
            else st

    ckpt = []
    st = state0
    for t in range(n_steps):
        if t % K == 0:
            ckpt.append(st)
        st, _ = step(at(st, t), model, cfg, advect=advect)
    fin = st
    n = fin.n
    det = np.linalg.det(fin.F)
    bars = (np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)),
            det[:, None, None] * _T(np.linalg.inv(fin.F)), np.zeros((n, 3, 3)))
    Gb = np.zeros((n, 3, 3))
    for s in range((n_steps - 1) // K, -1, -1):
        t0, t1 = s * K, min(n_steps, (s + 1) * K)
        caches, st = [], ckpt[s]
        for t in range(t0, t1):
            st, c = step(at(st, t), model, cfg, advect=advect)
            caches.append(c)
        for t in range(t1 - 1, t0 - 1, -1):
            if sched is None:
                bars = step_vjp(caches[t - t0], *bars, advect=advect)
            else:
                bars = step_vjp(caches[t - t0], *bars[:4], np.zeros_like(Gb), advect=advect)
                Gb = Gb + sched[t] * bars[4]
    if sched is None:
        Gb = bars[4]
    return float(np.sum(det)), float(np.sum(Gb * dG)), fin
