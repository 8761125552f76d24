"""Caller-side growth drivers and losses for the GPU trajectory.

The reference keeps the growth parameterisation in the caller (Fg = g I and
d/dg = sum tr Fg_bar, python/morphompm/integrate.py:31-38) and hard-wires the
loss to sum det F (:41-49).  These are the drivers and losses the workloads
need, as differentiable torch functions around ``morphompm_amd.autograd``:

  * ``ControlGridGrowth``: per-particle anisotropic Fg = I + diag(a, b, c),
    (a, b, c) trilinearly interpolated at the rest positions from a control
    grid of ctrl^3 x 3 values; its adjoint scatter is a fixed-order segment sum
    (particles sorted by control node once, float64 prefix sums), so gradients
    are bitwise reproducible;
  * ``position_misfit``: mean_p |x_p - x_target_p|^2;
  * ``tip_deflection``: mean y of the particles flagged as the tip;
  * ``sum_det``: the reference loss sum_p det F_p.
"""
import numpy as np
import torch


def _trilinear(pts, ctrl):
    u = np.clip(pts, 0.0, 1.0) * (ctrl - 1)
    i0 = np.minimum(np.floor(u).astype(np.int64), ctrl - 2)
    f = u - i0
    W = np.empty((len(pts), 8))
    idx = np.empty((len(pts), 8), dtype=np.int64)
    k = 0
    for a in range(2):
        for b in range(2):
            for c in range(2):
                W[:, k] = ((f[:, 0] if a else 1 - f[:, 0]) * (f[:, 1] if b else 1 - f[:, 1])
                           * (f[:, 2] if c else 1 - f[:, 2]))
                idx[:, k] = ((i0[:, 0] + a) * ctrl + (i0[:, 1] + b)) * ctrl + (i0[:, 2] + c)
                k += 1
    return W, idx


class _CtrlToDiag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, grid):
        ctx.grid = grid
        t = theta.reshape(-1, 3)
        return (grid.W.to(theta.dtype)[:, :, None] * t[grid.idx]).sum(1)      # (n, 3)

    @staticmethod
    def backward(ctx, gdiag):
        g = ctx.grid
        contrib = (g.W[:, :, None] * gdiag.to(torch.float64)[:, None, :]).reshape(-1, 3)
        # prefix sums along the contiguous (innermost) axis: a scan down the
        # 3-column axis of a tall array runs nearly serially on the device
        srt = contrib[g.order].T.contiguous()                                  # (3, 8 n)
        cs = torch.cat([srt.new_zeros(3, 1), torch.cumsum(srt, 1)], 1)
        out = torch.zeros(g.n_ctrl, 3, dtype=torch.float64, device=gdiag.device)
        out[g.nodes] = (cs[:, g.ends] - cs[:, g.starts]).T
        return out.reshape(g.shape).to(gdiag.dtype), None


class ControlGridGrowth:
    """Fg(theta) = I + diag(sum_k W_pk theta[idx_pk]) for a ctrl^3 x 3 theta."""

    def __init__(self, pts, ctrl, device="cuda"):
        W, idx = _trilinear(np.asarray(pts, dtype=np.float64), ctrl)
        self.ctrl = ctrl
        self.shape = (ctrl, ctrl, ctrl, 3)
        self.n_ctrl = ctrl ** 3
        self.W = torch.tensor(W, dtype=torch.float64, device=device)
        self.idx = torch.tensor(idx, device=device)
        flat = idx.reshape(-1)
        order = np.argsort(flat, kind="stable")
        nodes, starts, counts = np.unique(flat[order], return_index=True, return_counts=True)
        self.order = torch.tensor(order, device=device)
        self.nodes = torch.tensor(nodes, device=device)
        self.starts = torch.tensor(starts, device=device)
        self.ends = torch.tensor(starts + counts, device=device)

    def fg(self, theta):
        d = _CtrlToDiag.apply(theta, self)
        n = d.shape[0]
        eye = torch.eye(3, dtype=theta.dtype, device=theta.device).expand(n, 3, 3)
        return eye + torch.diag_embed(d)


def position_misfit(x, target):
    return ((x - target) ** 2).sum(1).mean()


def tip_deflection(x, tip_mask):
    return x[tip_mask, 1].mean()


def sum_det(F):
    return torch.linalg.det(F).sum()


def bilayer_curvature(x, n_bins=9):
    """Curvature of a bent strip from particle positions (n, 3), the way the
    reference measures it (tests/test_growth.cpp:266-295): the centreline is
    the mean y per x-bin (n_bins bins over the x extent), fitted by least
    squares with y = a xc^2 + b xc + c around the mean bin x; kappa = |2a|."""
    x = np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float64)
    lo, hi = x[:, 0].min(), x[:, 0].max()
    w = (hi - lo) / n_bins + 1e-9
    b = np.clip(((x[:, 0] - lo) / w).astype(np.int64), 0, n_bins - 1)
    cnt = np.bincount(b, minlength=n_bins)
    used = cnt > 0
    xb = np.bincount(b, x[:, 0], n_bins)[used] / cnt[used]
    yb = np.bincount(b, x[:, 1], n_bins)[used] / cnt[used]
    xc = xb - xb.mean()
    a = np.linalg.lstsq(np.stack([xc * xc, xc, np.ones_like(xc)], 1), yb, rcond=None)[0][0]
    return abs(2.0 * a)

