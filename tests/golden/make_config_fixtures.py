"""Write oracle results for the workloads at their real sizes as fixtures, so
the GPU tests compare against the FP64 oracle where it takes minutes to run.

Run in the build container (the oracle is this repository's own FP64 numpy
restatement, oracle/ref_vec.py; nothing from the reference is imported):

    python tests/golden/make_config_fixtures.py

Cases (every input is rebuilt by the tests from the same seeded builders in
morphompm_amd.scenes; the fixture holds the expected outputs):
  * batched inverse growth: four instances of the 256-instance batch (seed 0)
    -- the smallest, the median-sized neo-Hookean free swell and Hencky
    bilayer, and the largest -- at their true rates, 200 ramped steps with
    advection, checkpoint every 20: loss sum det F, its rate gradient, and
    x and F of 64 sampled particles;
  * control-grid shell (913,536 particles, 100^3, Hencky, trilinear
    anisotropic growth): x and F of 2,000 sampled particles after 3 steps;
  * bioink blob (2,097,152 particles, 128^3, Herschel-Bulkley): x and F of
    2,000 sampled particles after 2 steps.
Output: config_fixtures.npz (data only).
"""
import os
import sys
import time
from multiprocessing import get_context

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

N_SAMPLE_BATCH = 64
N_SAMPLE_SCENE = 2000
SHELL_STEPS = 3
BIOINK_STEPS = 2


def oracle_model(m):
    from oracle import ref_loop as RL
    if m.kind == "hb":
        return RL.HerschelBulkley(m.kappa, m.K, m.n, m.tau_y, m.m, m.eps)
    o = RL.NeoHookean(RL.Material()) if m.kind == "nh" else RL.Hencky(RL.Material())
    o.mu, o.lam = m.mu, m.lam
    return o


def oracle_cfg(cfg):
    from oracle import ref_loop as RL
    return RL.SimConfig(N=cfg.N, dx=cfg.dx, dt=cfg.dt, damping=cfg.damping, sp=cfg.sp,
                        material=RL.Material(cfg.material.E, cfg.material.nu,
                                             cfg.material.density), dims=cfg.dims)


def batch_indices(scenes):
    n = np.array([s.n for s in scenes])
    med = np.median(n)
    even = [i for i in range(len(n)) if i % 2 == 0]
    odd = [i for i in range(len(n)) if i % 2 == 1]
    return [int(np.argmin(n)), min(even, key=lambda i: abs(n[i] - med)),
            min(odd, key=lambda i: abs(n[i] - med)), int(np.argmax(n))]


def sample(n, k):
    return np.unique(np.linspace(0, n - 1, min(n, k)).astype(np.int64))


def batch_job(i):
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import ref_loop as RL, ref_vec as RV
    from morphompm_amd.scenes import batch_instances
    s = batch_instances(count=256, n_steps=200)[i]
    rate = s.extra["rate"]
    st0 = RL.ParticleState.rest(s.pts)
    st0.Fg = np.eye(3) + rate * s.dG
    t = time.time()
    L, g, fin = RV.det_loss_grad_checkpointed(st0, oracle_model(s.model), oracle_cfg(s.cfg),
                                              s.n_steps, s.dG, advect=True, sched=s.sched,
                                              K=20)
    idx = sample(s.n, N_SAMPLE_BATCH)
    print(f"batch instance {i}: n {s.n} {s.model.kind} L {L:.12e} g {g:.12e} "
          f"({time.time() - t:.0f} s)", flush=True)
    return i, s.n, rate, L, g, idx, fin.x[idx], fin.F[idx]


def scene_job(name):
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import ref_loop as RL, ref_vec as RV
    from morphompm_amd.scenes import bioink_blob, shell
    sc, steps = (shell(), SHELL_STEPS) if name == "shell" else (bioink_blob(), BIOINK_STEPS)
    st = RL.ParticleState.rest(sc.pts)
    st.Fg = sc.G.copy()
    t = time.time()
    fin, _ = RV.rollout(st, oracle_model(sc.model), oracle_cfg(sc.cfg), steps, advect=True,
                        keep=False)
    idx = sample(sc.n, N_SAMPLE_SCENE)
    print(f"{name}: n {sc.n}, {steps} steps ({time.time() - t:.0f} s)", flush=True)
    return name, sc.n, steps, idx, fin.x[idx], fin.F[idx]


def main():
    from morphompm_amd.scenes import batch_instances
    idx = batch_indices(batch_instances(count=256, n_steps=200))
    ctx = get_context("spawn")
    with ctx.Pool(len(idx) + 2) as pool:
        scenes = pool.map_async(scene_job, ["shell", "bioink"])
        res = pool.map(batch_job, idx)
        sres = scenes.get()
    out = {"batch_index": np.array([r[0] for r in res]), "batch_n": np.array([r[1] for r in res]),
           "batch_rate": np.array([r[2] for r in res]), "batch_L": np.array([r[3] for r in res]),
           "batch_g": np.array([r[4] for r in res]),
           "batch_sample": np.stack([np.pad(r[5], (0, N_SAMPLE_BATCH - len(r[5])),
                                            constant_values=-1) for r in res]),
           "batch_x": np.stack([np.pad(r[6], ((0, N_SAMPLE_BATCH - len(r[6])), (0, 0)))
                                for r in res]),
           "batch_F": np.stack([np.pad(r[7], ((0, N_SAMPLE_BATCH - len(r[7])), (0, 0), (0, 0)))
                                for r in res])}
    for name, n, steps, i, x, F in sres:
        out.update({f"{name}_n": n, f"{name}_steps": steps, f"{name}_sample": i,
                    f"{name}_x": x, f"{name}_F": F})
    np.savez_compressed(os.path.join(HERE, "config_fixtures.npz"), **out)
    print("wrote config_fixtures.npz", flush=True)


if __name__ == "__main__":
    main()
