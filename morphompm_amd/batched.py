"""Many growth-inversion instances in one launch sequence per iteration.

``BatchedInverse`` lays out instances of uneven particle counts in one
``BatchPlan`` and runs, per Gauss-Newton iteration, one forward trajectory and
one checkpointed adjoint over all of them: every kernel launch covers the
whole batch, and both phases are captured once as HIP graphs and replayed.

Growth parameterisation: instance i has a scalar rate theta_i and per-particle
direction dG_p, G_p = I + theta_i dG_p, ramped by the instance's schedule.
Loss (the reference's, python/morphompm/integrate.py:41-49): L_i = sum_p det F_p
at the final step, in float64.  dL_i/dtheta_i = sum_p <dG_p, G_bar_p>.
The update is the reference's damped Gauss-Newton (python/morphompm/diff.py:9-24)
applied per instance.
"""
import numpy as np
import torch

from .diff import gauss_newton_update
from .engine import BatchPlan, InstanceSpec, Trajectory, DEFAULT_TILE, ROW_F


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
                 use_graphs=True, checkpoint_every=None, keep_segments=1):
        """``tile`` / ``block``: the P2G tile shape and particle-kernel block
        size, or "auto" to pick them with the autotuner (cached per shape
        class).  ``keep_segments``: step-record stores the forward pass keeps
        (see ``engine.Trajectory``); more trade HBM for less recompute."""
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
        self.eye9 = torch.zeros(9, p.n_slots, dtype=torch.float64, device=p.device)
        self.eye9[[0, 4, 8]] = p.live.to(torch.float64)
        self.slot_inst = torch.from_numpy(p.slot_inst_np).to(p.device)
        self.G = p.new_mat(dtype)
        self.bar = p.new_dyn(dtype)
        self.use_graphs = use_graphs
        self._g_fwd = self._g_bwd = None
        self._final = None

    @property
    def particles(self):
        return sum(s.n for s in self.scenes)

    def set_theta(self, theta):
        th = torch.as_tensor(np.asarray(theta, dtype=np.float64), device=self.plan.device)
        G = self.eye9 + th[self.slot_inst][None, :] * self.dG
        self.G.copy_(G.to(self.dtype))

    # -- the two phases, each one launch sequence over the whole batch --------
    def _fwd(self):
        self._final = self.traj.forward(self.dyn0, self.G)

    def final_state(self):
        """The final particle state [24, n_slots] of the last forward run.  The
        adjoint's recompute reuses the segment store, so this reloads the last
        segment when needed (read it before the next forward)."""
        return self.traj.final()

    def _bwd(self):
        self.traj.backward(self.bar)

    def _capture(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up outside capture
            self._fwd()
            self._bwd()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self._g_fwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_fwd):
            self._fwd()
        self._seg_after_fwd = self.traj.store_state()
        self._g_bwd = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._g_bwd):
            self._bwd()
        self._seg_after_bwd = self.traj.store_state()

    def run_forward(self):
        if self.use_graphs:
            if self._g_fwd is None:
                self._capture()
            self._g_fwd.replay()
            # a replay leaves the segment store as the captured run did
            self.traj.set_store_state(self._seg_after_fwd)
        else:
            self._fwd()
        return self._final

    def run_backward(self):
        if self.use_graphs:
            if self._g_bwd is None:
                self._capture()
            self._g_bwd.replay()
            self.traj.set_store_state(self._seg_after_bwd)
        else:
            self._bwd()
        return self.traj.G_bar

    def loss_and_grad(self):
        """Per-instance (L_i, dL_i/dtheta_i) as float64 tensors."""
        p = self.plan
        fin = self.run_forward()
        Ff = fin[ROW_F:ROW_F + 9].to(torch.float64)
        det, cof = cofactor_rows(Ff)
        loss = p.per_instance_sum(det)
        self.bar.zero_()
        self.bar[ROW_F:ROW_F + 9] = (cof * p.live).to(self.dtype)
        Gb = self.run_backward()
        g = (Gb.to(torch.float64) * self.dG).sum(0)
        return loss, p.per_instance_sum(g)

    def infer(self, targets, theta0, iters=50, tol=1e-9, return_history=False):
        """Reference Gauss-Newton per instance; all instances are recomputed every
        iteration (one launch sequence), converged ones keep their theta."""
        theta = np.array(theta0, dtype=np.float64)
        targets = np.asarray(targets, dtype=np.float64)
        done = np.zeros(len(theta), dtype=bool)
        its = np.zeros(len(theta), dtype=np.int64)
        hist = [[t] for t in theta]
        for it in range(iters):
            self.set_theta(theta)
            L, dL = (t.cpu().numpy() for t in self.loss_and_grad())
            for i in np.nonzero(~done)[0]:
                its[i] = it + 1
                theta[i], done[i], appended = gauss_newton_update(theta[i], L[i], dL[i],
                                                                  targets[i], tol)
                if appended:
                    hist[i].append(theta[i])
            if done.all():
                break
        return (theta, its, hist) if return_history else (theta, its)
