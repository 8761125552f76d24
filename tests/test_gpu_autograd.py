"""torch.autograd.Function over the checkpointed GPU rollout."""
import numpy as np
import pytest
import torch

from oracle import ref_loop as RL
from oracle import ref_vec as RV

pytestmark = pytest.mark.gpu


def _inputs(pts, g, dtype):
    n = len(pts)
    dev = "cuda"
    x0 = torch.tensor(pts, dtype=dtype, device=dev)
    v0 = torch.zeros(n, 3, dtype=dtype, device=dev)
    C0 = torch.zeros(n, 3, 3, dtype=dtype, device=dev)
    F0 = torch.eye(3, dtype=dtype, device=dev).repeat(n, 1, 1)
    Fg = (g * torch.eye(3, dtype=dtype, device=dev)).repeat(n, 1, 1)
    return x0, v0, C0, F0, Fg


@pytest.mark.parametrize("kind", ["nh", "hencky"])
def test_autograd_matches_oracle_adjoint(kind):
    from morphompm_amd import autograd as AG, constitutive as K
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    m = K.NeoHookean(cfg.material) if kind == "nh" else K.Hencky(cfg.material)
    xs = [t.requires_grad_() for t in _inputs(RL.GATE_POINTS, 1.25, torch.float64)]
    x, v, C, F = AG.rollout(*xs, m, cfg, 5, advect=True, checkpoint_every=2)
    loss = (x ** 2).sum() + torch.linalg.det(F).sum() + (v * 0.3).sum()
    loss.backward()
    om = RL.NeoHookean(cfg.material) if kind == "nh" else RL.Hencky(cfg.material)
    st = RV.iso_growth_state(RL.GATE_POINTS, 1.25)
    fin, caches = RV.rollout(st, om, cfg, 5, advect=True)
    n = len(RL.GATE_POINTS)
    Fb = np.linalg.det(fin.F)[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
    bars = RV.rollout_vjp(caches, 2 * fin.x, 0.3 * np.ones((n, 3)), np.zeros((n, 3, 3)), Fb,
                          np.zeros((n, 3, 3)), advect=True)
    for t, ref in zip(xs, bars):
        g = t.grad.cpu().numpy()
        assert np.max(np.abs(g - ref)) <= 1e-10 * max(np.max(np.abs(ref)), 1e-30)


def test_autograd_gradcheck_fp64():
    """torch.autograd.gradcheck (central FD in FP64) through the GPU rollout,
    w.r.t. positions and growth."""
    from morphompm_amd import autograd as AG, constitutive as K
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    m = K.Hencky(cfg.material)
    x0, v0, C0, F0, Fg = _inputs(RL.GATE_POINTS[:3], 1.2, torch.float64)
    x0.requires_grad_()
    Fg.requires_grad_()

    def f(x0_, Fg_):
        x, v, C, F = AG.rollout(x0_, v0, C0, F0, Fg_, m, cfg, 3, advect=True)
        return torch.cat([x.reshape(-1), F.reshape(-1)])
    assert torch.autograd.gradcheck(f, (x0, Fg), eps=1e-7, atol=1e-6, rtol=1e-5)


def test_autograd_hb_params_fd():
    from morphompm_amd import autograd as AG, constitutive as K
    from morphompm_amd.config import SimConfig
    cfg = SimConfig()
    m = K.HerschelBulkley()
    ins = _inputs(RL.GATE_POINTS, 1.3, torch.float64)

    def L(theta):
        x, v, C, F = AG.rollout(*ins, m, cfg, 3, advect=True, hb_params=theta)
        return 0.5 * (x ** 2).sum() + torch.linalg.det(F).sum()
    th = torch.tensor([50.0, 1.0], dtype=torch.float64, device="cuda", requires_grad=True)
    L(th).backward()
    for k, h in ((0, 1e-4), (1, 1e-6)):
        e = torch.zeros(2, dtype=torch.float64, device="cuda")
        e[k] = h
        fd = (L(th.detach() + e) - L(th.detach() - e)).item() / (2 * h)
        assert abs(th.grad[k].item() - fd) <= 1e-5 * abs(fd) + 1e-12, (k, th.grad[k].item(), fd)
