"""Trajectories and their adjoint on the GPU (python/morphompm/integrate.py:15-52).

    rollout(state0, model, cfg, n_steps, advect=False)       -> (state, caches)
    rollout_vjp(caches, xf_b, vf_b, Cf_b, Ff_b, Fgf_b, advect=False)
    iso_growth_state(pts, g), iso_growth_grad(Fg_b), loss_detF(F)
    dloss_dg(pts, model, cfg, g, n_steps, advect)             -> (sum det F, d/dg)

Same signatures and defaults as the reference.  ``caches`` is an opaque
trajectory handle: instead of every step's cache it keeps a checkpoint every
``checkpoint_every`` steps (keyword; default: all steps in one segment) and
``rollout_vjp`` recomputes one segment at a time.  Keyword extensions:
``sched`` (per-step growth factors a_t with Fg_t = I + a_t (Fg - I); the
returned state's Fg is then the last step's, and the Fg cotangent returned is
that of the input Fg) and ``dtype``.
Loss and gradient sums accumulate in float64.
"""
import numpy as np
import torch

from ._convert import Like, to_dev
from .engine import Trajectory
from .state import ParticleState
from .transfer import plan_for


class RolloutHandle:
    """What ``rollout`` returns in place of the reference's list of caches."""

    def __init__(self, plan, traj, like):
        self.plan, self.traj, self.like = plan, traj, like

    def __len__(self):
        return self.traj.n_steps


def rollout(state0, model, cfg, n_steps, advect=False, *, checkpoint_every=None, sched=None,
            dtype=None):
    like = Like(state0.x, dtype)
    dt = like.dtype
    plan = plan_for(state0.n, cfg, model)
    traj = Trajectory(plan, dt, n_steps, advect, checkpoint_every, sched)
    dyn0 = plan.pack_dyn([(state0.x, state0.v, state0.C, state0.F)], dt)
    G = plan.pack_mat([state0.Fg], dt)
    fin = traj.forward(dyn0, G)
    x, v, C, F = plan.unpack_dyn(fin)
    Fg = plan.unpack_mat(G)
    if traj.sched is not None and traj.n_steps > 0:
        # the last step's growth tensor, I + a (G - I) (a == 1: G itself)
        a = float(traj.sched[traj.n_steps - 1, 0])
        if a != 1.0:
            eye = torch.eye(3, dtype=Fg.dtype, device=Fg.device)
            Fg = eye + a * (Fg - eye)
    st = ParticleState(like.out(x), like.out(v), like.out(C), like.out(F), like.out(Fg))
    return st, RolloutHandle(plan, traj, like)


def rollout_vjp(caches, xf_b, vf_b, Cf_b, Ff_b, Fgf_b, advect=False):
    h = caches
    if bool(advect) != h.traj.advect:
        raise ValueError("rollout_vjp: advect must match the rollout")
    p, dt = h.plan, h.traj.dtype
    bar = p.pack_dyn([(xf_b, vf_b, Cf_b, Ff_b)], dt, positions=False)
    Gb = p.pack_mat([Fgf_b], dt)
    bar0, G_bar, _ = h.traj.backward(bar, Gb)
    x, v, C, F = p.unpack_dyn(bar0, positions=False)
    o = h.like.out
    return o(x), o(v), o(C), o(F), o(p.unpack_mat(G_bar))


def iso_growth_state(pts, g):
    st = ParticleState.rest(pts)
    if isinstance(st.Fg, torch.Tensor):
        st.Fg[:] = g * torch.eye(3, dtype=st.Fg.dtype, device=st.Fg.device)
    else:
        st.Fg[:] = g * np.eye(3)
    return st


def iso_growth_grad(Fg_b):
    """d/dg for Fg = g I: sum_p tr(Fg_b[p]) (integrate.py:37-38), float64."""
    t = to_dev(Fg_b, torch.float64)
    return float(torch.diagonal(t, dim1=-2, dim2=-1).sum())


def loss_detF(F):
    """sum_p det F[p] (integrate.py:41-42), float64."""
    return float(torch.linalg.det(to_dev(F, torch.float64)).sum())


def detF_seed(F, dtype):
    """Cotangent of sum det F: det F F^-T, evaluated in float64."""
    Fd = F.to(torch.float64)
    return (torch.linalg.det(Fd)[:, None, None] * torch.linalg.inv(Fd).transpose(1, 2)).to(dtype)


def dloss_dg(pts, model, cfg, g, n_steps, advect, *, dtype=torch.float64, checkpoint_every=None):
    """d(sum det F_final)/dg through the composed GPU adjoint (integrate.py:45-52)."""
    n = len(pts)
    plan = plan_for(n, cfg, model)
    traj = Trajectory(plan, dtype, n_steps, advect, checkpoint_every)
    eye = torch.eye(3, dtype=dtype, device="cuda").expand(n, 3, 3)
    zeros3 = torch.zeros(n, 3, dtype=dtype, device="cuda")
    dyn0 = plan.pack_dyn([(to_dev(pts, dtype), zeros3, torch.zeros(n, 3, 3, dtype=dtype,
                                                                    device="cuda"), eye)], dtype)
    G = plan.pack_mat([g * eye], dtype)
    fin = traj.forward(dyn0, G)
    F = plan.unpack_dyn(fin)[3]
    loss = float(torch.linalg.det(F.to(torch.float64)).sum())
    bar = torch.zeros_like(fin)
    bar[15:24, plan.sl(0)] = detF_seed(F, dtype).reshape(n, 9).T
    _, G_bar, _ = traj.backward(bar)
    dg = float(torch.diagonal(plan.unpack_mat(G_bar).to(torch.float64), dim1=-2, dim2=-1).sum())
    return loss, dg
