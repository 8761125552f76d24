"""Differentiable GPU trajectory as a ``torch.autograd.Function``.

``rollout(x0, v0, C0, F0, Fg, model, cfg, n_steps, ...)`` returns the final
(x, v, C, F) of an n_steps MLS-MPM growth trajectory; its backward is the
hand-written reverse kernel set (python/morphompm/transfer.py:82-158 composed
as in python/morphompm/integrate.py:24-28), with checkpoints every
``checkpoint_every`` steps and recompute in between.  Gradients flow to x0, v0,
C0, F0, Fg and, for Herschel-Bulkley models, to ``hb_params`` = (tau_y, K).
"""
import torch

from .engine import ROW_C, ROW_F, ROW_V, ROW_X, BatchPlan, InstanceSpec, Trajectory


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dyn0, G, hb_params, traj, plan, prm_base):
        if hb_params is not None:
            prm = list(prm_base)
            prm[3], prm[1] = float(hb_params[0]), float(hb_params[1])   # tau_y, K
            plan.set_prm(0, prm)
        fin = traj.forward(dyn0, G).clone()
        ctx.traj = traj
        ctx.has_theta = hb_params is not None
        return fin

    @staticmethod
    def backward(ctx, bar):
        bar0, G_bar, theta = ctx.traj.backward(bar.contiguous(), want_theta=ctx.has_theta)
        th = theta[0].to(bar.dtype) if ctx.has_theta else None
        return bar0.clone(), G_bar.clone(), th, None, None, None


def _soa(t, n, k):
    return t.reshape(n, k).transpose(0, 1)


def rollout(x0, v0, C0, F0, Fg, model, cfg, n_steps, advect=False, *, checkpoint_every=None,
            sched=None, hb_params=None):
    """Final (x, v, C, F) after n_steps; differentiable.  Inputs are CUDA
    tensors (n,3)/(n,3,3) of one dtype (float32 or float64)."""
    n = x0.shape[0]
    dtype = x0.dtype
    # positions are stored relative to the cell nearest their mean (FP32
    # keeps more bits of the displacements); the plan is private to the call
    spec = InstanceSpec.from_cfg(n, cfg, model, x0.detach(), center=True)
    plan = BatchPlan([spec])
    traj = Trajectory(plan, dtype, n_steps, advect, checkpoint_every, sched)
    sl = plan.sl(0)
    pad = plan.n_slots
    off = plan.origin_offset(0)
    xr = (x0.to(torch.float64) - off).to(dtype)
    dyn0 = torch.cat([_soa(xr, n, 3), _soa(v0, n, 3), _soa(C0, n, 9), _soa(F0, n, 9)], 0)
    dyn0 = torch.nn.functional.pad(dyn0, (sl.start, pad - sl.stop))
    G = torch.nn.functional.pad(_soa(Fg, n, 9), (sl.start, pad - sl.stop))
    prm_base = plan.specs[0].prm
    if hb_params is not None and getattr(model, "kind", None) != "hb":
        raise ValueError("hb_params applies to HerschelBulkley models only")
    fin = _Rollout.apply(dyn0, G, hb_params, traj, plan, prm_base)
    f = fin[:, sl]
    x = (f[ROW_X:ROW_X + 3].T.to(torch.float64) + off).to(dtype)
    return (x, f[ROW_V:ROW_V + 3].T, f[ROW_C:ROW_C + 9].T.reshape(n, 3, 3),
            f[ROW_F:ROW_F + 9].T.reshape(n, 3, 3))
