"""Differentiable GPU trajectory as a ``torch.autograd.Function``.

``rollout(x0, v0, C0, F0, Fg, model, cfg, n_steps, ...)`` returns the final
(x, v, C, F) of an n_steps MLS-MPM growth trajectory; its backward is the
hand-written reverse kernel set (python/morphompm/transfer.py:82-158 composed
as in python/morphompm/integrate.py:24-28), with checkpoints every
``checkpoint_every`` steps and recompute in between.  Gradients flow to x0, v0,
C0, F0, Fg and, for Herschel-Bulkley models, to ``hb_params`` = (tau_y, K).
"""
import torch

from .engine import ROW_C, ROW_F, ROW_V, ROW_X, BatchPlan, DEFAULT_TILE, InstanceSpec, Trajectory


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


def _setup(x0, model, cfg, n_steps, advect, checkpoint_every, sched, keep_segments=1,
           tile=None, block=256, Fg=None):
    n = x0.shape[0]
    # positions are stored relative to the cell nearest their mean (FP32
    # keeps more bits of the displacements)
    spec = InstanceSpec.from_cfg(n, cfg, model, x0.detach(), center=True)
    if tile == "auto" or block == "auto":
        from .autotune import autotune
        G = Fg.detach() if Fg is not None else torch.eye(3, dtype=x0.dtype).expand(n, 3, 3)
        tile, block = autotune([spec], [x0.detach()], x0.dtype, advect, [G],
                               tiles=None if tile == "auto" else [tile or DEFAULT_TILE],
                               blocks=(64, 128, 256) if block == "auto" else (block,)).choice
    plan = BatchPlan([spec], tile=tile or DEFAULT_TILE, block=block)
    traj = Trajectory(plan, x0.dtype, n_steps, advect, checkpoint_every, sched,
                      keep_segments=keep_segments)
    return plan, traj, _cell_perm(plan, x0)


def _cell_perm(plan, x0):
    """Slot order of the particles: by bin of the rest positions (the caller's
    order comes back out of ``_unpack``).  Consecutive slots then hold nearby
    particles, so the stencil boxes that G2P and the reverse kernel stage in
    LDS per block of slots stay small whatever order the caller uses."""
    perm = torch.as_tensor(plan.cell_order([x0.detach()])[0], device=x0.device)  # sorted on x0's device
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(perm.numel(), device=x0.device)
    return perm, inv


def _pack(plan, order, x0, v0, C0, F0, Fg):
    perm = order[0]
    x0, v0, C0, F0, Fg = (t[perm] for t in (x0, v0, C0, F0, Fg))
    n = x0.shape[0]
    sl = plan.sl(0)
    pad = plan.n_slots
    xr = (x0.to(torch.float64) - plan.origin_offset(0)).to(x0.dtype)
    dyn0 = torch.cat([_soa(xr, n, 3), _soa(v0, n, 3), _soa(C0, n, 9), _soa(F0, n, 9)], 0)
    dyn0 = torch.nn.functional.pad(dyn0, (sl.start, pad - sl.stop))
    G = torch.nn.functional.pad(_soa(Fg, n, 9), (sl.start, pad - sl.stop))
    return dyn0, G


def _unpack(plan, order, fin, n, dtype):
    inv = order[1]
    f = fin[:, plan.sl(0)]
    x = (f[ROW_X:ROW_X + 3].T.to(torch.float64) + plan.origin_offset(0)).to(dtype)
    return (x[inv], f[ROW_V:ROW_V + 3].T[inv], f[ROW_C:ROW_C + 9].T[inv].reshape(n, 3, 3),
            f[ROW_F:ROW_F + 9].T[inv].reshape(n, 3, 3))


def _check_hb(model, hb_params):
    if hb_params is not None and getattr(model, "kind", None) != "hb":
        raise ValueError("hb_params applies to HerschelBulkley models only")


def rollout(x0, v0, C0, F0, Fg, model, cfg, n_steps, advect=False, *, checkpoint_every=None,
            sched=None, hb_params=None, keep_segments=1, tile=None, block=256):
    """Final (x, v, C, F) after n_steps; differentiable.  Inputs are CUDA
    tensors (n,3)/(n,3,3) of one dtype (float32 or float64).  The plan and
    its buffers are private to the call.  ``tile`` / ``block``: the P2G tile
    and particle block size (None: the default tile; "auto": the autotuner,
    cached per shape class); ``keep_segments``: see ``engine.Trajectory``."""
    _check_hb(model, hb_params)
    plan, traj, order = _setup(x0, model, cfg, n_steps, advect, checkpoint_every, sched,
                               keep_segments, tile, block, Fg)
    dyn0, G = _pack(plan, order, x0, v0, C0, F0, Fg)
    fin = _Rollout.apply(dyn0, G, hb_params, traj, plan, plan.specs[0].prm)
    return _unpack(plan, order, fin, x0.shape[0], x0.dtype)


class _GraphRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dyn0, G, hb_params, runner):
        fin = runner._forward(dyn0, G, hb_params)
        ctx.runner = runner
        return fin

    @staticmethod
    def backward(ctx, bar):
        bar0, G_bar, th = ctx.runner._backward(bar)
        return bar0, G_bar, th, None


class GraphRollout:
    """Repeated differentiable rollouts of one scene, for inversion loops: the
    same trajectory as ``rollout`` (bit for bit), but its forward pass and its
    checkpointed reverse pass are each captured once as a HIP graph and
    replayed on later calls, which removes the per-step launch work of the
    eager path.  Built from the rest positions x0 (the layout and tile
    alignment); each call takes the same arguments as ``rollout`` minus the
    fixed ones.  Inputs are copied into the runner's own buffers; outputs and
    gradients are fresh tensors.  One runner serves one call at a time.
    ``keep_segments``: step-record stores kept from the forward pass (see
    ``engine.Trajectory``); more trade HBM for less recompute.  ``tile`` /
    ``block`` as for ``rollout`` (``Fg``: a growth tensor the autotuner may
    time with)."""

    def __init__(self, x0, model, cfg, n_steps, advect=False, *, checkpoint_every=None,
                 sched=None, hb=False, keep_segments=1, tile=None, block=256, Fg=None):
        if hb and getattr(model, "kind", None) != "hb":
            raise ValueError("hb applies to HerschelBulkley models only")
        self.model, self.n, self.dtype, self.hb = model, x0.shape[0], x0.dtype, bool(hb)
        self.plan, self.traj, self.order = _setup(x0, model, cfg, n_steps, advect,
                                                  checkpoint_every, sched, keep_segments, tile,
                                                  block, Fg)
        p = self.plan
        self.prm_base = list(p.specs[0].prm)
        self.dyn_in = p.new_dyn(self.dtype)
        self.G_in = p.new_mat(self.dtype)
        self.bar_in = p.new_dyn(self.dtype)
        self._g_fwd = self._g_bwd = None

    def _set_theta(self, hb_params):
        if self.hb:
            prm = list(self.prm_base)
            prm[3], prm[1] = float(hb_params[0]), float(hb_params[1])   # tau_y, K
            self.plan.set_prm(0, prm)

    def _capture(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up outside capture
            self.traj.forward(self.dyn_in, self.G_in)
            self.traj.backward(self.bar_in, want_theta=self.hb)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self._g_fwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_fwd):
            self._fin = self.traj.forward(self.dyn_in, self.G_in)
        self._seg_after_fwd = self.traj.store_state()
        self._g_bwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_bwd):
            self._out = self.traj.backward(self.bar_in, want_theta=self.hb)
        self._seg_after_bwd = self.traj.store_state()

    def _forward(self, dyn0, G, hb_params):
        self.dyn_in.copy_(dyn0)
        self.G_in.copy_(G)
        self._set_theta(hb_params)
        if self._g_fwd is None:
            self._capture()
        self._g_fwd.replay()
        self.traj.set_store_state(self._seg_after_fwd)
        return self._fin.clone()

    def _backward(self, bar):
        self.bar_in.copy_(bar)
        self._g_bwd.replay()
        self.traj.set_store_state(self._seg_after_bwd)
        bar0, G_bar, theta = self._out
        th = theta[0].to(bar.dtype) if self.hb else None
        return bar0.clone(), G_bar.clone(), th

    def __call__(self, x0, v0, C0, F0, Fg, hb_params=None):
        if self.hb != (hb_params is not None):
            raise ValueError("hb_params must be given exactly when the runner was built with hb=True")
        if x0.shape[0] != self.n or x0.dtype != self.dtype:
            raise ValueError(f"runner built for {self.n} particles of {self.dtype}")
        dyn0, G = _pack(self.plan, self.order, x0, v0, C0, F0, Fg)
        fin = _GraphRollout.apply(dyn0, G, hb_params, self)
        return _unpack(self.plan, self.order, fin, self.n, self.dtype)
