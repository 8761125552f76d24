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


@pytest.mark.parametrize("kind, dtype", [("hencky", torch.float32), ("hb", torch.float64)])
def test_graph_rollout_replays_match_eager(kind, dtype):
    """GraphRollout (forward and reverse captured once as HIP graphs) against
    the eager rollout on the same inputs: outputs and every gradient bit for
    bit, over several calls with changed growth (and Herschel-Bulkley
    parameters), with eager rollouts run between the replays."""
    from morphompm_amd import autograd as AG, constitutive as K
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=6, ppc=8, seed=4)
    mat = sc.cfg.material
    m = (K.Hencky(mat) if kind == "hencky"
         else K.HerschelBulkley(kappa=mat.lam + 2 * mat.mu / 3, K=10.0, n=0.6))
    hb = kind == "hb"
    x0, v0, C0, F0, _ = _inputs(sc.pts, 1.0, dtype)
    runner = AG.GraphRollout(x0, m, sc.cfg, 12, advect=True, checkpoint_every=5, hb=hb)

    def run(fn, g, th):
        Fg = (g * torch.eye(3, dtype=dtype, device="cuda")).repeat(sc.n, 1, 1).requires_grad_()
        v = v0.clone().requires_grad_()
        thp = torch.tensor(th, dtype=dtype, device="cuda", requires_grad=True) if hb else None
        x, vn, C, F = fn(x0, v, C0, F0, Fg, thp)
        L = (x.double() ** 2).sum() + torch.linalg.det(F.double()).sum() + vn.double().sum()
        L.backward()
        return [t.detach() for t in (x, vn, C, F, Fg.grad, v.grad)] + ([thp.grad] if hb else [])

    def eager(x, v, C, F, Fg, th):
        return AG.rollout(x, v, C, F, Fg, m, sc.cfg, 12, advect=True, checkpoint_every=5,
                          hb_params=th)
    for rep, (g, th) in enumerate(((1.06, [50.0, 10.0]), (1.10, [40.0, 12.0]),
                                   (1.03, [60.0, 8.0]))):
        ref = run(eager, g, th)
        got = run(runner, g, th)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(ref, got)):
            assert torch.equal(a, b), (rep, i)


def test_rollout_tile_choice_fp64(tmp_path, monkeypatch):
    """rollout's tile/block choice (an explicit tile, or the autotuner's)
    changes only the summation order: FP64 results agree to 1e-11 of each
    output's largest entry."""
    from morphompm_amd import autograd as AG, constitutive as K
    from morphompm_amd.scenes import free_swell_cube
    monkeypatch.setenv("MORPHOMPM_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sc = free_swell_cube(cells=5, ppc=8, seed=1)
    m = K.Hencky(sc.cfg.material)
    x0, v0, C0, F0, Fg = _inputs(sc.pts, 1.08, torch.float64)
    ref = AG.rollout(x0, v0, C0, F0, Fg, m, sc.cfg, 6, advect=True, checkpoint_every=3)
    for kw in (dict(tile=(4, 8, 8), block=128), dict(tile="auto", block="auto")):
        got = AG.rollout(x0, v0, C0, F0, Fg, m, sc.cfg, 6, advect=True, checkpoint_every=3, **kw)
        for a, b in zip(ref, got):     # relative to each array's largest entry
            assert float((a - b).abs().max() / b.abs().max()) <= 1e-11, kw
