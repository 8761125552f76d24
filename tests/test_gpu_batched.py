"""Batched launches: many instances of uneven size in one plan, against the
FP64 oracle per instance and against single-instance runs."""
import numpy as np
import pytest
import torch

from oracle import ref_loop as RL
from oracle import ref_vec as RV

pytestmark = pytest.mark.gpu


def _oracle_model(m):
    if m.kind == "hb":
        return RL.HerschelBulkley(m.kappa, m.K, m.n, m.tau_y, m.m, m.eps)
    mat = RL.Material()
    o = RL.NeoHookean(mat) if m.kind == "nh" else RL.Hencky(mat)
    o.mu, o.lam = m.mu, m.lam
    return o


def _oracle_cfg(cfg):
    return RL.SimConfig(N=cfg.N, dx=cfg.dx, dt=cfg.dt, damping=cfg.damping, sp=cfg.sp,
                        material=RL.Material(cfg.material.E, cfg.material.nu,
                                             cfg.material.density), dims=cfg.dims)


def _small_batch(models=("nh", "hencky", "nh", "hencky"), n_steps=12, seed=1):
    from morphompm_amd import constitutive as K
    from morphompm_amd.scenes import batch_instances
    sc = batch_instances(count=len(models), seed=seed, n_lo=200, n_hi=900, n_steps=n_steps)
    for s, m in zip(sc, models):
        mat = s.cfg.material
        s.model = {"nh": K.NeoHookean(mat), "hencky": K.Hencky(mat),
                   "hb": K.HerschelBulkley(kappa=mat.lam + 2 * mat.mu / 3, K=10.0, n=0.6)}[m]
        s.checkpoint_every = 5
    return sc


def _run_batch(scenes, dtype, sched=True):
    from morphompm_amd.batched import BatchedInverse
    bi = BatchedInverse(scenes, dtype=dtype, use_graphs=False)
    bi.set_theta([s.extra["rate"] for s in scenes])
    L, g = bi.loss_and_grad()
    return bi, L.cpu().numpy(), g.cpu().numpy()


def _oracle_instance(s, rate):
    cfg = _oracle_cfg(s.cfg)
    m = _oracle_model(s.model)
    st0 = RL.ParticleState.rest(s.pts)
    st0.Fg = np.eye(3) + rate * s.dG
    fin, caches = RV.rollout(st0, m, cfg, s.n_steps, advect=s.advect, sched=s.sched)
    det = np.linalg.det(fin.F)
    n = s.n
    Fb = det[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
    z3, z33 = np.zeros((n, 3)), np.zeros((n, 3, 3))
    Gb = RV.rollout_vjp(caches, z3, z3, z33, Fb, z33, advect=s.advect, sched=s.sched)[4]
    return fin, det.sum(), float(np.sum(Gb * s.dG))


@pytest.mark.parametrize("models", [("nh", "hencky", "nh", "hencky"),
                                    ("hencky", "hencky", "hencky"),
                                    ("nh", "hb", "hencky")])
def test_batched_fp64_matches_oracle(models):
    sc = _small_batch(models)
    bi, L, g = _run_batch(sc, torch.float64)
    p = bi.plan
    for i, s in enumerate(sc):
        fin, Lo, go = _oracle_instance(s, s.extra["rate"])
        x, v, C, F = (t.cpu().numpy() for t in p.unpack_dyn(bi.final_state(), i, bi.perm))
        assert np.max(np.abs(F - fin.F)) <= 1e-10 * np.max(np.abs(fin.F)), i
        assert np.max(np.abs(x - fin.x)) <= 1e-12, i
        assert abs(L[i] - Lo) <= 1e-12 * abs(Lo), i
        assert abs(g[i] - go) <= 1e-9 * abs(go), (i, g[i], go)


def test_batched_without_advection_matches_oracle_fp64():
    """advect=False (positions fixed, transfer.py:75-77; x_bar = 0) over a
    whole batch: FP64 loss and rate gradient against the oracle, and the
    particles stay where they started."""
    sc = _small_batch(("nh", "hencky", "hb"), n_steps=10)
    for s in sc:
        s.advect = False
    bi, L, g = _run_batch(sc, torch.float64)
    assert bi.advect is False
    p = bi.plan
    for i, s in enumerate(sc):
        fin, Lo, go = _oracle_instance(s, s.extra["rate"])
        x, v, C, F = (t.cpu().numpy() for t in p.unpack_dyn(bi.final_state(), i, bi.perm))
        assert np.array_equal(x, s.pts) or np.max(np.abs(x - s.pts)) <= 1e-15, i
        assert np.max(np.abs(F - fin.F)) <= 1e-10 * np.max(np.abs(fin.F)), i
        assert abs(L[i] - Lo) <= 1e-12 * abs(Lo), i
        assert abs(g[i] - go) <= 1e-9 * abs(go), (i, g[i], go)


def test_every_tile_shape_matches_oracle_fp64():
    """Each built P2G tile shape (the autotuner's choices) gives the oracle's
    losses and gradients in FP64."""
    from morphompm_amd import _lib
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("nh", "hencky"), n_steps=8)
    ref = [_oracle_instance(s, s.extra["rate"]) for s in sc]
    for tile in _lib.tile_shapes():
        bi = BatchedInverse(sc, dtype=torch.float64, tile=tuple(tile), use_graphs=False)
        bi.set_theta([s.extra["rate"] for s in sc])
        L, g = (t.cpu().numpy() for t in bi.loss_and_grad())
        for i, (_, Lo, go) in enumerate(ref):
            assert abs(L[i] - Lo) <= 1e-12 * abs(Lo), (tile, i)
            assert abs(g[i] - go) <= 1e-9 * abs(go), (tile, i, g[i], go)


def test_batched_fp32_error_vs_fp64():
    sc = _small_batch(("nh", "hencky", "nh", "hencky"), n_steps=30)
    _, L32, g32 = _run_batch(sc, torch.float32)
    _, L64, g64 = _run_batch(sc, torch.float64)
    assert np.max(np.abs(L32 - L64) / np.abs(L64)) < 1e-5
    assert np.max(np.abs(g32 - g64) / np.abs(g64)) < 1e-5


def test_batched_equals_single_instance_bitwise():
    """Each instance in a batch gives exactly the bits it gives alone (same
    slot order and tile alignment): instances never interact."""
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory
    sc = _small_batch(("nh", "hencky", "nh"), n_steps=8)
    bi, _, _ = _run_batch(sc, torch.float32)
    for i, s in enumerate(sc):
        p = BatchPlan([InstanceSpec.from_cfg(s.n, s.cfg, s.model, s.pts, center=True)])
        perm = p.cell_order([s.pts])
        tr = Trajectory(p, torch.float32, s.n_steps, True, 5, s.sched)
        dyn0 = p.pack_dyn([(s.pts, np.zeros((s.n, 3)), np.zeros((s.n, 3, 3)),
                            np.repeat(np.eye(3)[None], s.n, 0))], torch.float32, perm)
        G = p.pack_mat([np.eye(3) + s.extra["rate"] * s.dG], torch.float32, perm)
        fin = tr.forward(dyn0, G)
        a = [t.cpu() for t in p.unpack_dyn(fin, 0, perm)]
        b = [t.cpu() for t in bi.plan.unpack_dyn(bi.final_state(), i, bi.perm)]
        for u, w in zip(a, b):
            assert torch.equal(u, w), i


def test_graph_replays_match_eager():
    """Captured forward/adjoint graphs replayed several times, with eager work
    in between, reproduce the eager result every time."""
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("nh", "hencky"), n_steps=10)
    eager = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    graph = BatchedInverse(sc, dtype=torch.float32, use_graphs=True)
    rates = np.array([s.extra["rate"] for s in sc])
    for rep in range(3):
        th = rates * (1.0 + 0.1 * rep)
        eager.set_theta(th)
        graph.set_theta(th)
        Le, ge = eager.loss_and_grad()
        Lg, gg = graph.loss_and_grad()
        torch.cuda.synchronize()
        assert torch.equal(Le, Lg) and torch.equal(ge, gg), rep
        # the final state after replays (the segment store was left holding
        # the first segment by the replayed adjoint; an eager final_state()
        # before the replays must not make it look current)
        assert torch.equal(eager.final_state(), graph.final_state()), rep


def test_final_state_after_eager_call_and_replays():
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("nh", "hencky"), n_steps=12)    # 3 segments of 5, 5, 2
    graph = BatchedInverse(sc, dtype=torch.float32, use_graphs=True)
    graph.set_theta([s.extra["rate"] for s in sc])
    graph.loss_and_grad()
    ref = graph.final_state().clone()                   # eager rerun of the last segment
    graph.loss_and_grad()                               # replays: the store ends on segment 0
    assert torch.equal(graph.final_state(), ref)


def test_hb_theta_batched_fd_fp64():
    """Herschel-Bulkley (tau_y, K) cotangents from the batched adjoint against
    central differences of the FP64 GPU forward."""
    from morphompm_amd import constitutive as K
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory, ROW_F
    from morphompm_amd.batched import cofactor_rows
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=4, origin=6, N=16, n_steps=6, dt=1e-4)
    mat = sc.cfg.material

    def run(ty, Kc, want_grad):
        m = K.HerschelBulkley(kappa=mat.lam + 2 * mat.mu / 3, K=Kc, n=0.6, tau_y=ty)
        p = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, m, sc.pts)])
        tr = Trajectory(p, torch.float64, sc.n_steps, True, 3, sc.sched)
        dyn0 = p.pack_dyn([(sc.pts, np.zeros((sc.n, 3)), np.zeros((sc.n, 3, 3)),
                            np.repeat(np.eye(3)[None], sc.n, 0))], torch.float64)
        fin = tr.forward(dyn0, p.pack_mat([sc.G], torch.float64))
        det, cof = cofactor_rows(fin[ROW_F:ROW_F + 9])
        L = float((det * p.live).sum())
        if not want_grad:
            return L
        bar = torch.zeros_like(fin)
        bar[ROW_F:ROW_F + 9] = cof * p.live
        _, _, th = tr.backward(bar, want_theta=True)
        return L, th[0].cpu().numpy()
    _, th = run(50.0, 10.0, True)
    for k, (dty, dK) in enumerate([(1e-3, 0.0), (0.0, 1e-4)]):
        fd = (run(50.0 + dty, 10.0 + dK, False) - run(50.0 - dty, 10.0 - dK, False)) / (2 * (dty + dK))
        assert abs(th[k] - fd) <= 1e-6 * abs(fd) + 1e-14, (k, th[k], fd)


def test_batched_infer_recovers_rates():
    """Per-instance Gauss-Newton (diff.py:9-24 semantics) over a mixed batch in
    FP64: every instance converges to its true rate from 0.7x it, and its
    iterates equal those of a host loop over the batch's own (L, L')."""
    from morphompm_amd.batched import BatchedInverse
    from morphompm_amd.diff import gauss_newton_update
    sc = _small_batch(("nh", "hencky", "nh", "hencky"), n_steps=8)
    rates = np.array([s.extra["rate"] for s in sc])
    bi = BatchedInverse(sc, dtype=torch.float64, use_graphs=False)
    bi.set_theta(rates)
    target = bi.loss_and_grad()[0].cpu().numpy()
    theta, its, hist = bi.infer(target, 0.7 * rates, iters=30, tol=1e-9, return_history=True)
    assert np.all(its < 30), its
    np.testing.assert_allclose(theta, rates, rtol=1e-6)
    # the same iteration on the host for instance 1, one evaluation at a time
    th, k = 0.7 * rates, 1
    seq = [th[k]]
    for _ in range(30):
        bi.set_theta(th)
        L, dL = (t.cpu().numpy() for t in bi.loss_and_grad())
        g, done, app = gauss_newton_update(th[k], L[k], dL[k], target[k], 1e-9)
        if done:
            break
        th = th.copy()
        th[k] = g
        if app:
            seq.append(g)
    np.testing.assert_allclose(hist[k], seq, rtol=0, atol=1e-12)


def test_kernel_timer_eager_and_under_capture():
    """mpm_kernel_timer: eager launches of the armed kernel are timed; an
    armed timer does not disturb stream capture (captured launches are not
    timed) and the replays match the eager result."""
    import ctypes
    from morphompm_amd import _lib
    from morphompm_amd.batched import BatchedInverse
    L = _lib.lib()
    sc = _small_batch(("nh", "hencky"), n_steps=6)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(64)]
    for e in evs:
        e.record()
    torch.cuda.synchronize()
    h = (ctypes.c_void_p * 64)(*[e.cuda_event for e in evs])
    eager = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    eager.set_theta([s.extra["rate"] for s in sc])
    assert L.mpm_kernel_timer(4, h, 32) == 0          # G2P
    try:
        Le, ge = eager.loss_and_grad()
        torch.cuda.synchronize()
        n_eager = L.mpm_kernel_timer_count()
        assert n_eager > 0
        assert all(evs[2 * i].elapsed_time(evs[2 * i + 1]) > 0 for i in range(n_eager))
        graph = BatchedInverse(sc, dtype=torch.float32, use_graphs=True)
        graph.set_theta([s.extra["rate"] for s in sc])
        L.mpm_kernel_timer(4, h, 32)
        Lg, gg = graph.loss_and_grad()                 # warm-up (eager, timed), capture, replay
        torch.cuda.synchronize()
        assert torch.equal(Le, Lg) and torch.equal(ge, gg)
        assert L.mpm_kernel_timer_count() == n_eager   # only the eager warm-up was timed
    finally:
        L.mpm_kernel_timer(-1, None, 0)


def test_kernel_timer_is_per_host_thread():
    """An armed timer times the launches of the thread that armed it only:
    an iteration run from another host thread records nothing in it."""
    import ctypes
    import threading
    from morphompm_amd import _lib
    from morphompm_amd.batched import BatchedInverse
    L = _lib.lib()
    sc = _small_batch(("nh", "hencky"), n_steps=6)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(64)]
    for e in evs:
        e.record()
    torch.cuda.synchronize()
    h = (ctypes.c_void_p * 64)(*[e.cuda_event for e in evs])
    bi = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    bi.set_theta([s.extra["rate"] for s in sc])
    bi.loss_and_grad()                       # first launches outside the timed part
    torch.cuda.synchronize()
    assert L.mpm_kernel_timer(4, h, 32) == 0
    try:
        seen = {}

        def other():
            bi.loss_and_grad()
            torch.cuda.synchronize()
            seen["count"] = L.mpm_kernel_timer_count()
        t = threading.Thread(target=other)
        t.start()
        t.join()
        assert seen["count"] == 0 and L.mpm_kernel_timer_count() == 0
        bi.loss_and_grad()
        torch.cuda.synchronize()
        assert L.mpm_kernel_timer_count() > 0
    finally:
        L.mpm_kernel_timer(-1, None, 0)


def test_kept_segments_under_graphs_match_eager():
    """Every segment kept (no recompute) with the forward and reverse passes
    captured as HIP graphs: losses, gradients and final states equal the
    eager recompute path bit for bit over repeated replays."""
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("nh", "hencky"), n_steps=12)    # 3 segments of 5, 5, 2
    eager = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    kept = BatchedInverse(sc, dtype=torch.float32, use_graphs=True, keep_segments=3)
    assert kept.traj.M == 3
    rates = np.array([s.extra["rate"] for s in sc])
    for rep in range(3):
        th = rates * (1.0 + 0.05 * rep)
        eager.set_theta(th)
        kept.set_theta(th)
        Le, ge = eager.loss_and_grad()
        Lk, gk = kept.loss_and_grad()
        torch.cuda.synchronize()
        assert torch.equal(Le, Lk) and torch.equal(ge, gk), rep
        assert torch.equal(eager.final_state(), kept.final_state()), rep


def test_iteration_graph_matches_eager_device_loop():
    """A whole Gauss-Newton iteration (growth, forward, loss, reverse, gradient,
    update on the device) captured as one HIP graph: every replay, with eager
    work and synchronisations between replays, leaves the same rates,
    convergence flags, iteration counts and histories as the eager device
    loop, bit for bit; the graph is captured without advancing the state."""
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("nh", "hencky", "nh"), n_steps=10)
    rates = np.array([s.extra["rate"] for s in sc])
    eager = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    graph = BatchedInverse(sc, dtype=torch.float32, use_graphs=True)
    eager.set_theta(rates)
    target = eager.loss_and_grad()[0].cpu().numpy()
    for bi in (eager, graph):
        bi.start_inversion(target, 0.6 * rates, tol=1e-9, hist_cap=8)
    graph.prepare_iteration()
    torch.cuda.synchronize()
    assert torch.equal(graph.theta.cpu(), torch.tensor(0.6 * rates))
    assert int(graph.its.sum()) == 0
    for it in range(5):
        eager.gn_iterate(1)
        graph.gn_iterate(1)
        torch.cuda.synchronize()
        _ = eager.loss.sum().item()                   # eager work between replays
        for a, b in ((eager.theta, graph.theta), (eager.done, graph.done),
                     (eager.its, graph.its), (eager.hist, graph.hist),
                     (eager.hist_len, graph.hist_len), (eager.loss, graph.loss),
                     (eager.grad, graph.grad)):
            assert torch.equal(a, b), it
    # the iterates approach the true rates
    assert np.max(np.abs(graph.theta.cpu().numpy() - rates)) < 1e-3 * np.max(rates)


def test_device_update_matches_host_rule():
    """mpm_gauss_newton against diff.gauss_newton_update on every branch:
    converged, nudge (|L'| < 1e-12, L' = inf, L' = nan, both signs of the
    residual), a capped step either way, an uncapped step, and a NaN loss."""
    import ctypes
    from morphompm_amd import _lib
    from morphompm_amd._lib import ptr
    from morphompm_amd.diff import gauss_newton_update
    L = np.array([1.0, 2.0, 0.5, 2.0, 2.0, 10.0, -10.0, 1.3, np.nan, 1.0 + 1e-10])
    dL = np.array([2.0, 1e-13, 1e-13, np.inf, np.nan, 1.0, 1.0, 3.0, 2.0, 5.0])
    tgt = np.ones_like(L)
    th0 = np.linspace(0.5, 1.5, len(L))
    dev = lambda a, t=torch.float64: torch.tensor(a, dtype=t, device="cuda")
    th, done, its = dev(th0), dev(np.zeros(len(L)), torch.int32), dev(np.zeros(len(L)), torch.int32)
    hist = torch.zeros(len(L), 3, dtype=torch.float64, device="cuda")
    hist[:, 0] = th
    hl = torch.ones(len(L), dtype=torch.int32, device="cuda")
    dLoss, dGrad, dTgt = dev(L), dev(dL), dev(tgt)    # held until the kernel has run
    rc = _lib.lib().mpm_gauss_newton(len(L), ptr(dLoss), ptr(dGrad), ptr(dTgt), 1e-9,
                                     ptr(th), ptr(done), ptr(its), ptr(hist), 3, ptr(hl),
                                     _lib.stream_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    for i in range(len(L)):
        g, d, app = gauss_newton_update(th0[i], L[i], dL[i], tgt[i], 1e-9)
        assert th[i].item() == g and bool(done[i].item()) == d and its[i].item() == 1, i
        assert hl[i].item() == (2 if app else 1), i
        if app:
            assert hist[i, 1].item() == g, i


def test_multi_iteration_graph_matches_eager_device_loop():
    """Several Gauss-Newton iterations captured in one HIP graph
    (iterations_per_graph = 3): seven iterations (two replays of the
    3-iteration graph, then one single-iteration replay) leave the eager
    device loop's rates, flags, counts and histories, bit for bit."""
    from morphompm_amd.batched import BatchedInverse
    sc = _small_batch(("hencky", "nh"), n_steps=8)
    rates = np.array([s.extra["rate"] for s in sc])
    eager = BatchedInverse(sc, dtype=torch.float32, use_graphs=False)
    multi = BatchedInverse(sc, dtype=torch.float32, use_graphs=True, iterations_per_graph=3)
    eager.set_theta(rates)
    target = eager.loss_and_grad()[0].cpu().numpy()
    for bi in (eager, multi):
        bi.start_inversion(target, 0.5 * rates, tol=1e-9, hist_cap=9)
        bi.gn_iterate(7)
    torch.cuda.synchronize()
    for a, b in ((eager.theta, multi.theta), (eager.done, multi.done), (eager.its, multi.its),
                 (eager.hist, multi.hist), (eager.hist_len, multi.hist_len)):
        assert torch.equal(a, b)
    assert int(multi.its.max()) >= 3
