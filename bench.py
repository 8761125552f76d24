"""Forward + adjoint particle-steps/s of the batched growth inversion on one MI355X.

Workload (default): 256 independent instances, particle counts log-uniform in
[1k, 32k] (seeded), growth rates U(0.02, 0.2): neo-Hookean free-swell cubes
(even instances) and Hencky bilayer strips (odd instances), 200 steps each,
adjoint with checkpoints every 20 steps, FP32.  One bench step = one
Gauss-Newton iteration of the batched inversion: growth tensors from the
rates, the forward trajectory, the loss, the checkpointed adjoint (recompute
included) and the gradient of every instance, and the per-instance update --
all on the device, captured once as one HIP graph and replayed (the timed
region has no host round trip).  After the timed iterations the inversion
continues to 50 iterations and the bench reports how well each instance's rate
is recovered.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Besides the metric it reports the dominant kernel's
HBM roofline fraction against the bytes the reference data model moves (timed
with HIP events over one eager iteration after the timed region), the plain
single-instance GPU rate (same kernels, one instance at a time, no graph),
FP32 error vs an FP64 GPU run, and the FP64 numpy oracle timed on the host
cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("fwd+adjoint particle-steps/s, batched FP32 vs plain GPU, grad rel.err≤1e-5")
HBM_PEAK_GBS = 8000.0

# Kernel ids of the library's timer (include/morphompm_hip.h) with each
# kernel's compulsory bytes per launch (FP32): n particles, a active grid
# nodes, nb bins.  Per particle: records of 4-byte words read + written.
def g2p_bytes_per_particle(n_steps, K):
    """Average algorithmic bytes per particle of one G2P launch over an
    iteration (forward pass + the reverse pass's recompute of all but the last
    segment), from what each launch reads and writes: x, F, node v in; x, F
    out, plus C and v where the state is a record or a checkpoint; with a next
    step in the chain also G in and the next payload (64 B) and key out."""
    nseg = -(-n_steps // K)
    seg_len = [min(K, n_steps - s * K) for s in range(nseg)]
    total, launches = 0, 0

    def launch(out, nxt):
        rows = {"full": 96, "record": 84, "next": 48}[out]
        return 48 + rows + (36 + 68 if nxt else 0)
    for s, L in enumerate(seg_len):                       # forward pass
        keep = s == nseg - 1
        for k in range(L):
            last = k == L - 1
            out = "full" if last else ("record" if keep else "next")
            total += launch(out, not last)
            launches += 1
    for s, L in enumerate(seg_len[:-1]):                  # recompute of the other segments
        for k in range(L - 1):                            # (the last step runs no G2P)
            total += launch("record", True)
            launches += 1
    return total / launches


def adj_particle_bytes_per_particle(n_steps, K):
    """Average algorithmic bytes per particle of one per-particle reverse
    launch over an iteration (P2G-adjoint gather and constitutive VJP): rank,
    the forward payload (64 B), the adjoint payload (64 B), xn_b, G, C, F,
    C_new, Fn_b and G_b in (380 B); x_b, v_b, F_b and G_b out (96 B).  A
    launch that also prepares the previous step's adjoint payload reads x, F
    of that step and its rank (52 B) and writes the 64 B payload; the first
    step of a segment writes the C_b rows (36 B) instead: 592 and 512 B."""
    nseg = -(-n_steps // K)
    seg_len = [min(K, n_steps - s * K) for s in range(nseg)]
    total = sum((L - 1) * 592 + 512 for L in seg_len)
    return total / n_steps


# Bytes the reference data model (full 3x3 Fg, python/morphompm/state.py:10-16)
# moves per particle-step in FP32, independent of this implementation:
#   forward particles: read x, v, C, F, Fg (33 words), write x, v, C, F (24)  228 B
#   forward grid, per active node: P2G write m + mom (4), update read 4 /
#     write 3, G2P read 3                                                   56 B
#   reverse particles: read the input state (33), the next step's C (9), the
#     output cotangents x, v, C, F (24) and Fg_bar (9); write the input
#     cotangents (24) and Fg_bar (9)                                        432 B
#   reverse grid, per active node: v + m (4), v_bar write / read (3 + 3),
#     mom_bar + m_bar write / read (4 + 4)                                  72 B
# The per-particle reverse kernel (P2G-adjoint gather, constitutive VJP and
# the previous step's G2P-adjoint payload) does all of the reverse particle
# work, so its roofline is priced at 432 B per particle.
REF_FWD_PARTICLE, REF_FWD_NODE = 228, 56
REF_ADJ_PARTICLE, REF_ADJ_NODE = 432, 72


def reference_bytes_per_particle_step(r):
    """Whole forward + adjoint step, r = active nodes per particle (checkpoint
    recompute not counted)."""
    return REF_FWD_PARTICLE + REF_ADJ_PARTICLE + (REF_FWD_NODE + REF_ADJ_NODE) * r


# Implementation bytes (what each kernel of this build reads and writes),
# per launch: n particles, a active grid nodes, nb bins.
KERNELS = {
    0: ("sort", lambda n, a, nb: 40 * n + 12 * nb),         # key in/out, start, placed, rank,
                                                             # sorted; counts, cursors, starts
    1: ("prep_fwd", lambda n, a, nb: 200 * n),              # x v C F G in, payload + key out
                                                             # (first step of each chain only)
    2: ("p2g_fwd_tile", lambda n, a, nb: 68 * n + 16 * a),  # sorted + payload in, partials out
    3: ("grid_fwd", lambda n, a, nb: 32 * a),               # partials in, (m, v) out
    4: ("g2p", lambda n, a, nb: G2P_BYTES * n + 16 * a),    # see g2p_bytes_per_particle
    5: ("prep_adj", lambda n, a, nb: 208 * n),              # x F + 4 cotangents in, payload out
    6: ("p2g_adj_tile", lambda n, a, nb: 64 * n + 16 * a),
    7: ("grid_adj", lambda n, a, nb: 48 * a),               # partials + grid in, adjoint grid out
    8: ("adj_particle", lambda n, a, nb: ADJ_PARTICLE_BYTES * n + 32 * a),  # see
                                                             # adj_particle_bytes_per_particle
}
G2P_BYTES = 248.0   # set from the trajectory shape in main()
ADJ_PARTICLE_BYTES = 580.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--sim-steps", type=int, default=200)
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--plain-instances", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-keep-all", action="store_true",
                    help="skip the reference measurement with every segment's step records "
                         "kept in HBM (no recompute)")
    ap.add_argument("--tile", default="auto", help='P2G tile "x,y,z" or "auto" (autotuner)')
    ap.add_argument("--block", default="auto", help='particle-kernel block size or "auto"')
    ap.add_argument("--pmc-pass", action="store_true",
                    help="for rocprofv3 --pmc: a device copy of --copy-mb MiB (counter "
                         "calibration), then one eager forward + adjoint iteration; no timing")
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--no-others", action="store_true",
                    help="skip the single-scene workloads (bilayer strip, bioink blob, shell)")
    ap.add_argument("--inversion-iters", type=int, default=50,
                    help="Gauss-Newton iterations of the full inversion report (the timed "
                         "ones included); 0 reports none (profiling runs)")
    return ap.parse_args()


# ----------------------------------------------------------------- CPU leg
def _cpu_worker(args):
    """One single-thread process: FP64 numpy oracle forward + adjoint of one
    instance for a truncated number of steps."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import numpy as np
    from oracle import ref_vec as RV
    from oracle import ref_loop as RL
    pts, G, dx, dt, steps, sched, kind = args
    mat = RL.Material(E=1.0e3, nu=0.3, density=1.0)
    hi = np.ceil(pts.max(0) / dx).astype(int) + 4
    cfg = RL.SimConfig(N=int(hi.max()), dx=dx, dt=dt, sp=dx / 2, material=mat,
                       dims=tuple(int(h) for h in hi))
    st0 = RL.ParticleState.rest(pts)
    st0.Fg = G.copy()
    model = RL.NeoHookean(mat) if kind == "nh" else RL.Hencky(mat)
    t0 = time.perf_counter()
    fin, caches = RV.rollout(st0, model, cfg, steps, advect=True, sched=sched)
    n = len(pts)
    Fb = np.linalg.det(fin.F)[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
    RV.rollout_vjp(caches, np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), Fb,
                   np.zeros((n, 3, 3)), advect=True, sched=sched)
    return n * steps, time.perf_counter() - t0


def _cpu_loop_worker(args):
    """The loop-faithful restatement (the reference's per-particle x 27-node
    Python loops) for one forward + adjoint step of one instance, single
    thread."""
    import numpy as np
    from oracle import ref_loop as RL
    pts, G, dx, dt, steps, sched, kind = args
    mat = RL.Material(E=1.0e3, nu=0.3, density=1.0)
    hi = np.ceil(pts.max(0) / dx).astype(int) + 4
    cfg = RL.SimConfig(N=int(hi.max()), dx=dx, dt=dt, sp=dx / 2, material=mat,
                       dims=tuple(int(h) for h in hi))
    st0 = RL.ParticleState.rest(pts)
    st0.Fg = G.copy()
    n = len(pts)
    model = RL.NeoHookean(mat) if kind == "nh" else RL.Hencky(mat)
    t0 = time.perf_counter()
    fin, cache = RL.step(st0, model, cfg, advect=True)
    Fb = np.linalg.det(fin.F)[:, None, None] * np.swapaxes(np.linalg.inv(fin.F), 1, 2)
    RL.step_vjp(cache, np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), Fb,
                np.zeros((n, 3, 3)), advect=True)
    return n, time.perf_counter() - t0


def _cpu_model():
    """The host CPU's model name, as /proc/cpuinfo reports it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scenes, budget_s):
    """FP64 vectorised numpy oracle on the host: P single-thread processes over
    independent instances, each run for a truncated step count."""
    import multiprocessing as mp
    P = max(1, os.cpu_count() or 1)
    steps = 4
    jobs = []
    for s in scenes[: 4 * P]:
        jobs.append((s.pts, s.G, s.cfg.dx, s.cfg.dt, steps,
                     s.sched[:steps] if s.sched is not None else None, s.model.kind))
    # calibrate the number of jobs to the time budget with one job
    n0, t0 = _cpu_worker(jobs[0])
    per_ps = t0 / n0
    total_ps = sum(len(j[0]) * steps for j in jobs)
    while len(jobs) > P and total_ps * per_ps / P > budget_s:
        jobs.pop()
        total_ps = sum(len(j[0]) * steps for j in jobs)
    ctx = mp.get_context("spawn")
    t = time.perf_counter()
    with ctx.Pool(P) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t
    done = sum(r[0] for r in res)
    small = min(jobs, key=lambda j: len(j[0]))
    n_loop, t_loop = _cpu_loop_worker(small)
    return {"value": done / wall, "unit": "particle-steps/s", "cores": P, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "loop_faithful_single_thread_value": n_loop / t_loop,
            "loop_faithful_sample": f"1 fwd+adjoint step of a {n_loop}-particle instance "
                                    f"in {t_loop:.1f} s (the reference's loop structure)",
            "sample": f"{len(jobs)} of the instances, {steps} fwd+adjoint steps each "
                      f"(FP64 vectorised numpy oracle, 1 thread per process, {P} processes, "
                      f"{done} particle-steps in {wall:.1f} s)",
            "single_thread_value": n0 / t0}


def other_workloads():
    """Forward + adjoint particle-steps/s of the single-scene workloads at full
    size, FP32, through autograd.GraphRollout (captured once, then replayed;
    the replayed call is timed, inputs resident on the device): the bilayer
    strip (65,536 particles, 500 steps, Hencky), the bioink blob (2,097,152,
    300 steps, Herschel-Bulkley, gradient w.r.t. (tau_y, K)) and the
    control-grid shell (~0.91 M, 300 steps, Hencky).  The bioink adjoint
    checkpoints every 25 steps and recomputes in between, as that workload
    specifies; the bilayer and the shell specify no checkpointing, so their
    forward pass keeps every step record that fits in HBM
    (keep_segments="auto") and the adjoint recomputes only the rest."""
    import numpy as np
    import torch
    from morphompm_amd.autograd import GraphRollout
    from morphompm_amd.scenes import bilayer_strip, bioink_blob, shell
    out = {}
    for name, build, K, keep in (("bilayer_strip_65k_x500", bilayer_strip, 50, "auto"),
                                 ("bioink_2.1M_x300", bioink_blob, 25, 1),
                                 ("shell_0.9M_x300", shell, 25, "auto")):
        sc = build()
        n, dt = sc.n, torch.float32
        x0 = torch.tensor(sc.pts, dtype=dt, device="cuda")
        z3 = torch.zeros(n, 3, dtype=dt, device="cuda")
        z33 = torch.zeros(n, 3, 3, dtype=dt, device="cuda")
        F0 = torch.eye(3, dtype=dt, device="cuda").repeat(n, 1, 1)
        G = torch.tensor(sc.G, dtype=dt, device="cuda")
        hb = sc.model.kind == "hb"
        run = GraphRollout(x0, sc.model, sc.cfg, sc.n_steps, True, checkpoint_every=K, hb=hb,
                           keep_segments=keep)
        th = torch.tensor([sc.model.tau_y, sc.model.K], dtype=dt, device="cuda") if hb else None
        ms = []
        for rep in range(2):
            Gr = G.clone().requires_grad_(True)
            thr = th.clone().requires_grad_(True) if hb else None
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            x, v, C, F = run(x0, z3, z33, F0, Gr, thr)
            L = (x.double() ** 2).sum() + torch.linalg.det(F.double()).sum()
            L.backward()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        out[name] = {"particles": n, "steps": sc.n_steps, "checkpoint_every": K,
                     "segments_kept": run.traj.M,
                     "segments": -(-sc.n_steps // K),
                     "model": sc.model.kind, "ms_fwd_adj": ms[-1],
                     "particle_steps_per_s": n * sc.n_steps / (ms[-1] / 1e3),
                     "timed": "second call (graph replay), forward + adjoint + loss"}
        print(f"[bench] {name}: {out[name]['particle_steps_per_s']:.3g} particle-steps/s "
              f"({run.traj.M} of {out[name]['segments']} segments kept)",
              file=sys.stderr, flush=True)
        del run, x0, G, F0, z3, z33
        torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "requested_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from morphompm_amd import _lib
    from morphompm_amd.batched import BatchedInverse, cofactor_rows
    from morphompm_amd.engine import BatchPlan, InstanceSpec, Trajectory, ROW_F
    from morphompm_amd.scenes import batch_instances
    import ctypes

    _lib.require_gpu()
    tile = "auto" if a.tile == "auto" else tuple(int(t) for t in a.tile.split(","))
    block = "auto" if a.block == "auto" else int(a.block)
    scenes = batch_instances(count=a.instances, n_steps=a.sim_steps)
    for s in scenes:
        s.checkpoint_every = a.checkpoint_every
    n_total = sum(s.n for s in scenes)
    global G2P_BYTES, ADJ_PARTICLE_BYTES
    G2P_BYTES = g2p_bytes_per_particle(a.sim_steps, a.checkpoint_every)
    ADJ_PARTICLE_BYTES = adj_particle_bytes_per_particle(a.sim_steps, a.checkpoint_every)
    rates = np.array([s.extra["rate"] for s in scenes])
    models = sorted({s.model.kind for s in scenes})
    print(f"[bench] {len(scenes)} instances ({', '.join(models)}), {n_total} particles, "
          f"{a.sim_steps} steps, K={a.checkpoint_every}", file=sys.stderr, flush=True)

    if a.pmc_pass:
        # counters are read per kernel name: no tuning dispatches in this run
        from morphompm_amd.engine import DEFAULT_TILE
        tile = DEFAULT_TILE if tile == "auto" else tile
        block = 256 if block == "auto" else block
        src = torch.ones(a.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        bi = BatchedInverse(scenes, dtype=torch.float32, tile=tile, block=block,
                            use_graphs=False)
        bi.set_theta(rates)
        bi.loss_and_grad()
        torch.cuda.synchronize()
        print(json.dumps({"pmc_pass": True, "copy_bytes": a.copy_mb << 20,
                          "particles": n_total}))
        return
    t_tune = time.perf_counter()
    # the timed iterations are one graph launch
    bi = BatchedInverse(scenes, dtype=torch.float32, tile=tile, block=block, use_graphs=True,
                        iterations_per_graph=a.steps)
    tune = None
    if bi.tune is not None:
        tune = {"shape_class": bi.tune.key, "cached": bi.tune.cached,
                "ms_per_candidate_trajectory": bi.tune.ms,
                "setup_s": round(time.perf_counter() - t_tune, 2)}
        print(f"[bench] autotuner: tile {bi.plan.tile} block {bi.plan.block} "
              f"({'cached' if bi.tune.cached else 'tuned'})", file=sys.stderr, flush=True)
    tile, block = bi.plan.tile, bi.plan.block
    # targets: the loss at the true rates (the observed shapes)
    bi.set_theta(rates)
    target, _ = bi.loss_and_grad()
    # the inversion starts every instance at the rates' prior mean
    theta0 = np.full(len(scenes), 0.5 * (0.02 + 0.2))
    n_iter = max(a.inversion_iters, a.warmup + a.steps)
    bi.start_inversion(target, theta0, tol=1e-9, hist_cap=1)

    t_cap = time.perf_counter()
    bi.prepare_iteration()           # captures the iteration graphs
    torch.cuda.synchronize()
    capture_s = time.perf_counter() - t_cap
    bi.gn_iterate(a.warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    bi.gn_iterate(a.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_per = ms_total / a.steps
    ps_per_iter = n_total * a.sim_steps
    value = ps_per_iter * a.steps / (ms_total / 1e3)
    print(f"[bench] {ms_per:.2f} ms per iteration, {value:.4g} particle-steps/s "
          f"(wall {wall:.2f} s)", file=sys.stderr, flush=True)

    # the rest of the inversion, untimed: every instance's recovered rate
    if a.inversion_iters > 0:
        bi.gn_iterate(n_iter - a.warmup - a.steps)
    th = bi.theta.cpu().numpy()
    its = bi.its.cpu().numpy()
    done = bi.done.cpu().numpy().astype(bool)
    err = np.abs(th - rates)
    inversion = None
    if a.inversion_iters > 0:
        inversion = {"iterations": n_iter, "theta0": float(theta0[0]), "tol": 1e-9,
                     "converged": int(done.sum()), "instances": len(scenes),
                     "recovered_within_1e-4": int((err < 1e-4).sum()),
                     "max_abs_err": float(err.max()), "median_abs_err": float(np.median(err)),
                     "median_iterations": float(np.median(its)),
                     "max_iterations": int(its.max())}
    if inversion:
        print(f"[bench] inversion: {inversion['recovered_within_1e-4']}/{len(scenes)} rates "
              f"within 1e-4 after {n_iter} iterations (max err {err.max():.2e})",
              file=sys.stderr, flush=True)

    # Per-kernel timing: one more iteration per kernel, run eagerly on the
    # same stream with HIP events around each launch of that kernel (event
    # records inside captured graphs are refused on this stack).
    L = _lib.lib()
    cap = 2 * a.sim_steps
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * cap)]
    for e in evs:
        e.record()
    torch.cuda.synchronize()
    handles = (ctypes.c_void_p * (2 * cap))(*[e.cuda_event for e in evs])
    bi.set_theta(th)
    ktab = {}
    for kid, (kname, _) in KERNELS.items():
        L.mpm_kernel_timer(kid, handles, cap)
        bi._fwd()
        bi._loss()
        bi._bwd()
        cnt = L.mpm_kernel_timer_count()
        L.mpm_kernel_timer(-1, None, 0)
        torch.cuda.synchronize()
        d = [evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(cnt)]
        # the timer is armed for exactly one iteration: the sum is its share
        ktab[kname] = {"id": kid, "avg_ms": float(np.mean(d)) if d else 0.0,
                       "launches_timed": cnt, "ms_per_iteration": float(np.sum(d))}
    # active nodes: one more step of the final state into a zeroed grid record
    # (step records are written only where particles reach)
    p = bi.plan
    g0, sb0, pl0, d1 = p.new_grid(torch.float32), p.new_sort(), p.new_payload(torch.float32), p.new_dyn(torch.float32)
    p.step_fwd(bi.final_state(), bi.G, None, d1, g0, sb0, pl0, True)
    active = float(int((g0[:, 0] > 1e-12).sum()))
    del g0, sb0, pl0, d1
    # occupied bins of this step's sort: n_active units (last word of a sort record)
    occ_bins = float(np.mean([int(bi.traj.seg_sort[k][-1]) for k in range(bi.traj.seg_sort.shape[0])])
                     ) * int(np.prod(bi.plan.tile))
    for kname, e in ktab.items():
        kid = e["id"]
        e["implementation_bytes_per_launch"] = KERNELS[kid][1](n_total, active, occ_bins)
        e["implementation_GBs"] = (e["implementation_bytes_per_launch"] / (e["avg_ms"] / 1e3) / 1e9
                                   if e["avg_ms"] > 0 else 0.0)
    # the dominant single kernel (the sort entry is a 7-launch sequence)
    dom = max((k for k in ktab if k != "sort"), key=lambda k: ktab[k]["ms_per_iteration"])
    d = ktab[dom]
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_batched.txt")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith("traffic_bytes_per_launch"):
                traffic = float(line.split()[1])
    if dom == "adj_particle":
        ref_bytes = REF_ADJ_PARTICLE * n_total
        ref_model = ("reverse particle work of the reference data model: 432 B per particle "
                     "(read input state, next-step C, output cotangents, Fg_bar; write input "
                     "cotangents, Fg_bar)")
    else:
        ref_bytes = d["implementation_bytes_per_launch"]
        ref_model = "implementation bytes (no reference-model share for this kernel)"
    achieved = ref_bytes / (d["avg_ms"] / 1e3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": ref_bytes, "byte_model": ref_model,
            "traffic_over_algorithmic": traffic / ref_bytes if traffic else None,
            "implementation_bytes_per_launch": d["implementation_bytes_per_launch"],
            "avg_launch_ms": d["avg_ms"], "launches_timed": d["launches_timed"],
            "particles_per_launch": n_total,
            "timer": "hip events, eager iteration after the timed region"}

    # whole-iteration bytes of the reference data model (fwd + adjoint, per
    # particle-step; the checkpoint recompute is not counted)
    r = active / n_total
    b_alg = reference_bytes_per_particle_step(r)
    whole = {"B_alg_per_particle_step": b_alg, "active_nodes_per_particle": r,
             "achieved_GBs": b_alg * value / 1e9,
             "frac": b_alg * value / 1e9 / HBM_PEAK_GBS}

    # For reference (not the metric): the same iteration with the forward
    # pass keeping every segment's step records in HBM, so the reverse pass
    # recomputes nothing (checkpoints are still written every K steps).
    keep_all = None
    if not a.no_keep_all:
        nseg = -(-a.sim_steps // a.checkpoint_every)
        bk = BatchedInverse(scenes, dtype=torch.float32, tile=tile, block=block,
                            use_graphs=True, keep_segments=nseg)
        st = bk.traj._stores
        hbm = sum(t.numel() * t.element_size() for store in st for t in store)
        bk.start_inversion(target, theta0, tol=1e-9)
        bk.gn_iterate(1)                           # capture + warm-up
        torch.cuda.synchronize()
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record()
        bk.gn_iterate(a.steps)
        k1.record()
        torch.cuda.synchronize()
        ms_k = k0.elapsed_time(k1) / a.steps
        keep_all = {"value": ps_per_iter / (ms_k / 1e3), "ms_per_step": ms_k,
                    "keep_segments": nseg, "segment_stores_GB": hbm / 1e9,
                    "note": "reference only: no checkpoint recompute in the reverse pass"}
        del bk, st
        torch.cuda.empty_cache()
        print(f"[bench] every segment kept: {ms_k:.2f} ms per iteration "
              f"({hbm / 1e9:.0f} GB of step records)", file=sys.stderr, flush=True)

    # plain single-instance path: same kernels, one instance at a time, no graph
    plain = None
    if a.plain_instances > 0:
        idx = np.argsort([s.n for s in scenes])[len(scenes) // 2 - a.plain_instances // 2:][:a.plain_instances]
        tot_ps, tot_t = 0, 0.0
        for i in idx:
            s = scenes[i]
            p = BatchPlan([InstanceSpec.from_cfg(s.n, s.cfg, s.model, s.pts)], tile=tile,
                          block=block)
            tr = Trajectory(p, torch.float32, s.n_steps, True, a.checkpoint_every, s.sched)
            dyn0 = p.pack_dyn([(s.pts, np.zeros((s.n, 3)), np.zeros((s.n, 3, 3)),
                                np.repeat(np.eye(3)[None], s.n, 0))], torch.float32)
            G = p.pack_mat([np.eye(3) + th[i] * s.dG], torch.float32)
            for rep in range(2):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                fin = tr.forward(dyn0, G)
                bar = torch.zeros_like(fin)
                F = fin[ROW_F:ROW_F + 9].to(torch.float64)
                bar[ROW_F:ROW_F + 9] = (cofactor_rows(F)[1] * p.live).float()
                tr.backward(bar)
                torch.cuda.synchronize()
                dtt = time.perf_counter() - t1
            tot_ps += s.n * s.n_steps
            tot_t += dtt
        plain = {"per_instance_value": tot_ps / tot_t, "instances_sampled": int(len(idx)),
                 "median_n": int(np.median([scenes[i].n for i in idx])),
                 "models_sampled": [scenes[i].model.kind for i in idx]}

    # FP32 vs FP64 on the same batch (one forward + adjoint)
    acc = None
    if not a.no_fp64_check:
        bi.set_theta(th)
        L32, g32 = (t.cpu().numpy() for t in bi.loss_and_grad())
        F32 = bi.final_state()[ROW_F:ROW_F + 9].double().clone()
        del bi
        torch.cuda.empty_cache()
        b64 = BatchedInverse(scenes, dtype=torch.float64, tile=tile, block=block,
                             use_graphs=False)
        b64.set_theta(th)
        L64, g64 = (t.cpu().numpy() for t in b64.loss_and_grad())
        F64 = b64.final_state()[ROW_F:ROW_F + 9]
        live = b64.plan.live
        acc = {"forward_max_abs_dF": float((F32 - F64)[:, live].abs().max()),
               "loss_max_rel_err": float(np.max(np.abs(L32 - L64) / np.abs(L64))),
               "grad_max_rel_err": float(np.max(np.abs(g32 - g64) / np.abs(g64))),
               "grad_median_rel_err": float(np.median(np.abs(g32 - g64) / np.abs(g64)))}
        del b64
        torch.cuda.empty_cache()

    others = None if a.no_others else other_workloads()

    cpu = None
    if not a.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scenes, a.cpu_seconds)

    out = {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "batched inverse growth (Gauss-Newton iteration = growth, "
                                  "200-step forward, loss, checkpointed adjoint, gradient and "
                                  "update of every instance, on the device)",
                      "instances": len(scenes), "particles": n_total,
                      "models": "neo-Hookean free swell (even), Hencky bilayer (odd)",
                      "sim_steps": a.sim_steps, "checkpoint_every": a.checkpoint_every,
                      "tile": list(tile), "block": block,
                      "hip_graphs": f"the {a.steps} timed iterations are one graph launch "
                                    f"(captured in {capture_s:.1f} s)"},
           "autotune": tune,
           "roofline": roof, "whole_step_roofline": whole, "kernels": ktab,
           "cpu_baseline": cpu,
           "plain_gpu": plain,
           "no_recompute": keep_all,
           "batched_over_plain": (value / plain["per_instance_value"]) if plain else None,
           "batched_over_cpu": (value / cpu["value"]) if cpu else None,
           "fp32_vs_fp64": acc,
           "inversion": inversion,
           "single_scene_workloads": others}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
