"""Many growth-inversion instances in one launch sequence per iteration.

``BatchedInverse`` lays out instances of uneven particle counts in one
``BatchPlan`` and runs, per Gauss-Newton iteration, one forward trajectory and
one checkpointed adjoint over all of them: every kernel launch covers the
whole batch.  The caller-side steps run on the device too (HIP kernels of the
library: growth tensors from the rates, the loss and its seed cotangent, the
gradient reduction and the update), so one whole iteration is captured once as
a HIP graph and replayed with no host round trip.

Growth parameterisation: instance i has a scalar rate theta_i and per-particle
direction dG_p, G_p = I + theta_i dG_p, ramped by the instance's schedule.
Loss (the reference's, python/morphompm/integrate.py:41-49): L_i = sum_p det F_p
at the final step, in float64.  dL_i/dtheta_i = sum_p <dG_p, G_bar_p>.
The update is the reference's damped Gauss-Newton (python/morphompm/diff.py:9-24)
applied per instance.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import SLOT_ALIGN, check, ptr, stream_ptr, suffix
from .engine import BatchPlan, InstanceSpec, Trajectory, DEFAULT_TILE


def _f64(v, device):
    """Per-instance values (numpy, list or tensor) as a float64 device tensor."""
    if isinstance(v, torch.Tensor):
        return v.detach().to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(v, dtype=np.float64), device=device)


def cofactor_rows(F):
    """F rows [9, n] (row-major 3x3 per column) -> (det F [n], cof F rows [9, n]),
    cof F = det F F^-T, the cotangent of det F."""
    f = F
    c = torch.empty_like(f)
    c[0] = f[4] * f[8] - f[5] * f[7]
    c[1] = f[5] * f[6] - f[3] * f[8]
    c[2] = f[3] * f[7] - f[4] * f[6]
    c[3] = f[2] * f[7] - f[1] * f[8]
    c[4] = f[0] * f[8] - f[2] * f[6]
    c[5] = f[1] * f[6] - f[0] * f[7]
    c[6] = f[1] * f[5] - f[2] * f[4]
    c[7] = f[2] * f[3] - f[0] * f[5]
    c[8] = f[0] * f[4] - f[1] * f[3]
    det = f[0] * c[0] + f[1] * c[1] + f[2] * c[2]
    return det, c


class BatchedInverse:
    def __init__(self, scenes, dtype=torch.float32, tile=DEFAULT_TILE, block=256,
                 use_graphs=True, checkpoint_every=None, keep_segments=1,
                 iterations_per_graph=1):
        """``tile`` / ``block``: the P2G tile shape and particle-kernel block
        size, or "auto" to pick them with the autotuner (cached per shape
        class).  ``keep_segments``: step-record stores the forward pass keeps
        (see ``engine.Trajectory``); more trade HBM for less recompute.
        ``iterations_per_graph``: Gauss-Newton iterations captured in one HIP
        graph (``gn_iterate`` replays it for whole multiples and runs any
        remainder as single-iteration replays)."""
        self.scenes = scenes
        steps = {s.n_steps for s in scenes}
        if len(steps) != 1:
            raise ValueError("all instances of a batch take the same number of steps")
        self.n_steps = steps.pop()
        adv = {bool(s.advect) for s in scenes}
        if len(adv) != 1:
            raise ValueError("all instances of a batch share the advect flag")
        self.advect = adv.pop()
        self.dtype = dtype
        specs = [InstanceSpec.from_cfg(s.n, s.cfg, s.model, s.pts, center=True) for s in scenes]
        self.tune = None
        if tile == "auto" or block == "auto":
            from .autotune import autotune
            G_list = [np.eye(3) + s.extra.get("rate", 0.05) * s.dG for s in scenes]
            self.tune = autotune(specs, [s.pts for s in scenes], dtype, self.advect, G_list,
                                 tiles=None if tile == "auto" else [tile],
                                 blocks=(64, 128, 256) if block == "auto" else (block,))
            tile, block = self.tune.choice
        self.plan = p = BatchPlan(specs, tile=tile, block=block)
        K = checkpoint_every or scenes[0].checkpoint_every
        sched = np.stack([s.sched if s.sched is not None else np.ones(self.n_steps)
                          for s in scenes], 1)
        self.traj = Trajectory(p, dtype, self.n_steps, self.advect, K,
                               None if np.all(sched == 1.0) else sched,
                               keep_segments=keep_segments)
        n = [s.n for s in scenes]
        self.perm = p.cell_order([s.pts for s in scenes])
        self.dyn0 = p.pack_dyn([(s.pts, np.zeros((k, 3)), np.zeros((k, 3, 3)),
                                 np.repeat(np.eye(3)[None], k, 0)) for s, k in zip(scenes, n)],
                               dtype, self.perm)
        self.dG = p.pack_mat([s.dG for s in scenes], torch.float64, self.perm)
        # the growth kernel writes the trajectory's growth tensor in place
        self.G = self.traj.G
        self.bar = p.new_dyn(dtype)
        f64 = dict(dtype=torch.float64, device=p.device)
        i32 = dict(dtype=torch.int32, device=p.device)
        ni = p.n_inst
        self.theta = torch.zeros(ni, **f64)
        self.loss = torch.zeros(ni, **f64)
        self.grad = torch.zeros(ni, **f64)
        self.chunk_acc = torch.zeros(p.n_slots // SLOT_ALIGN, **f64)
        # Gauss-Newton state (start_inversion)
        self.target = torch.zeros(ni, **f64)
        self.done = torch.zeros(ni, **i32)
        self.its = torch.zeros(ni, **i32)
        self.hist = torch.zeros(ni, 1, **f64)
        self.hist_len = torch.zeros(ni, **i32)
        self.tol = 1e-9
        self.use_graphs = use_graphs
        self.iterations_per_graph = max(1, int(iterations_per_graph))
        self._g_fwd = self._g_bwd = None
        self._g_iter = {}        # iterations per graph -> captured graph
        self._g_iter_key = None
        self._final = None

    @property
    def particles(self):
        return sum(s.n for s in self.scenes)

    # -- device steps ---------------------------------------------------------
    def _call(self, name, *args):
        f = getattr(_lib.lib(), f"{name}_{suffix(self.dtype)}")
        check(f(ctypes.byref(self.plan.c), *args, stream_ptr()), name)

    def _growth(self):
        self._call("mpm_growth_from_rates", ptr(self.theta), ptr(self.dG), ptr(self.G))

    def _fwd(self):
        self._final = self.traj.forward(self.dyn0)

    def _loss(self):
        self._call("mpm_loss_det", ptr(self._final), ptr(self.bar), ptr(self.chunk_acc),
                   ptr(self.loss))

    def _bwd(self):
        self.traj.backward(self.bar)

    def _grad(self):
        self._call("mpm_growth_grad", ptr(self.traj.G_bar), ptr(self.dG), ptr(self.chunk_acc),
                   ptr(self.grad))

    def _update(self, state=None):
        theta, done, its, hist, hist_len = state or (self.theta, self.done, self.its, self.hist,
                                                     self.hist_len)
        check(_lib.lib().mpm_gauss_newton(self.plan.n_inst, ptr(self.loss), ptr(self.grad),
                                          ptr(self.target), float(self.tol), ptr(theta),
                                          ptr(done), ptr(its), ptr(hist), hist.shape[1],
                                          ptr(hist_len), stream_ptr()), "gauss_newton")

    def set_theta(self, theta):
        """Per-instance rates -> the growth tensors of the next forward pass."""
        self.theta.copy_(_f64(theta, self.plan.device))
        self._growth()

    def final_state(self):
        """The final particle state [24, n_slots] of the last forward run.  The
        adjoint's recompute reuses the segment store, so this reloads the last
        segment when needed (read it before the next forward)."""
        return self.traj.final()

    # -- graphs ---------------------------------------------------------------
    def _warm(self, fn):
        """Runs fn once on a side stream outside any capture (first launches
        load kernels and allocate), then waits for it."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()

    def _capture(self):
        def phases():
            self._fwd()
            self._loss()
            self._bwd()
            self._grad()
        self._warm(phases)
        self._g_fwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_fwd):
            self._fwd()
            self._loss()
        self._seg_after_fwd = self.traj.store_state()
        self._g_bwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_bwd):
            self._bwd()
            self._grad()
        self._seg_after_bwd = self.traj.store_state()

    def _capture_iteration(self, k):
        """k whole Gauss-Newton iterations as one graph: growth, forward,
        loss, reverse, gradient, update, k times.  The warm-up runs the update
        on a copy of the inversion state, so capturing does not advance it."""
        def phases(state):
            self._growth()
            self._fwd()
            self._loss()
            self._bwd()
            self._grad()
            self._update(state)
        scratch = tuple(t.clone() for t in (self.theta, self.done, self.its, self.hist,
                                            self.hist_len))
        theta0 = self.theta.clone()
        self._warm(lambda: phases(scratch))
        self.theta.copy_(theta0)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(k):
                phases(None)
        self._g_iter[k] = g
        self._seg_after_iter = self.traj.store_state()

    def run_forward(self):
        """Forward pass and loss of the current growth tensors."""
        if self.use_graphs:
            if self._g_fwd is None:
                self._capture()
            self._g_fwd.replay()
            # a replay leaves the segment store as the captured run did
            self.traj.set_store_state(self._seg_after_fwd)
        else:
            self._fwd()
            self._loss()
        return self._final

    def run_backward(self):
        """Adjoint of the last forward pass and the per-instance gradient."""
        if self.use_graphs:
            if self._g_bwd is None:
                self._capture()
            self._g_bwd.replay()
            self.traj.set_store_state(self._seg_after_bwd)
        else:
            self._bwd()
            self._grad()
        return self.traj.G_bar

    def loss_and_grad(self):
        """Per-instance (L_i, dL_i/dtheta_i) as float64 device tensors."""
        self.run_forward()
        self.run_backward()
        return self.loss.clone(), self.grad.clone()

    # -- inversion ------------------------------------------------------------
    def start_inversion(self, targets, theta0, tol=1e-9, hist_cap=1):
        """Resets the Gauss-Newton state: targets, initial rates (also the first
        history entry), nothing converged, no iterations."""
        self.target.copy_(_f64(targets, self.plan.device))
        self.theta.copy_(_f64(theta0, self.plan.device))
        self.tol = float(tol)
        if self.hist.shape[1] < hist_cap:
            self.hist = torch.zeros(self.plan.n_inst, int(hist_cap), dtype=torch.float64,
                                    device=self.plan.device)
        self.hist.zero_()
        self.hist[:, 0] = self.theta
        self.hist_len.fill_(1)
        self.done.zero_()
        self.its.zero_()

    def _graph(self, k):
        key = (self.tol, self.hist.data_ptr(), self.hist.shape[1])
        if self._g_iter_key != key:          # the update's arguments changed
            self._g_iter = {}
            self._g_iter_key = key
        if k not in self._g_iter:
            self._capture_iteration(k)
        return self._g_iter[k]

    def prepare_iteration(self):
        """Captures the iteration graphs now (when use_graphs), so that later
        gn_iterate calls only replay them."""
        if self.use_graphs:
            self._graph(self.iterations_per_graph)
            if self.iterations_per_graph > 1:
                self._graph(1)

    def gn_iterate(self, n=1):
        """n Gauss-Newton iterations of every unconverged instance, on the
        device (graph replays when use_graphs: n // iterations_per_graph
        replays of the multi-iteration graph, then single-iteration ones);
        no host synchronisation."""
        if not self.use_graphs:
            for _ in range(n):
                self._growth()
                self._fwd()
                self._loss()
                self._bwd()
                self._grad()
                self._update()
            return
        k = self.iterations_per_graph
        for _ in range(n // k):
            self._graph(k).replay()
            self.traj.set_store_state(self._seg_after_iter)
        for _ in range(n % k):
            self._graph(1).replay()
            self.traj.set_store_state(self._seg_after_iter)

    def infer(self, targets, theta0, iters=50, tol=1e-9, return_history=False):
        """Reference Gauss-Newton per instance (diff.py:9-24): all instances are
        recomputed every iteration (one launch sequence), converged ones keep
        their theta; stops when every instance has converged.  Returns (theta,
        iterations[, history]) as numpy."""
        self.start_inversion(targets, theta0, tol, hist_cap=iters + 1)
        for _ in range(iters):
            self.gn_iterate(1)
            if bool(self.done.all()):
                break
        theta = self.theta.cpu().numpy()
        its = self.its.cpu().numpy().astype(np.int64)
        if not return_history:
            return theta, its
        h, hl = self.hist.cpu().numpy(), self.hist_len.cpu().numpy()
        return theta, its, [list(h[i, :hl[i]]) for i in range(len(hl))]
