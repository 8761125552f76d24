"""bench.py's host side: the one-GPU guard and the byte models its roofline
figures are priced with (no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["requested_gpus"] == 2 and "one GPU" in out["error"]


def test_reference_byte_model_counts_its_words():
    """The reference data model's bytes per particle-step, recounted from the
    words each phase moves (FP32; full 3x3 Fg, python/morphompm/state.py:10-16):
    x 3, v 3, C 9, F 9, Fg 9 = 33 words of state."""
    state, dyn = 33, 24                     # state with Fg; x, v, C, F
    assert bench.REF_FWD_PARTICLE == 4 * (state + dyn)
    assert bench.REF_ADJ_PARTICLE == 4 * (state + 9 + dyn + 9 + dyn + 9)
    assert bench.REF_FWD_NODE == 4 * (4 + 4 + 3 + 3)
    assert bench.REF_ADJ_NODE == 4 * (4 + 3 + 3 + 4 + 4)
    assert bench.reference_bytes_per_particle_step(0.0) == 660
    assert bench.reference_bytes_per_particle_step(1.0) == 660 + 128
    assert bench.reference_bytes_per_particle_step(0.31) == pytest.approx(699.68)


def test_adjoint_kernel_byte_model_over_segments():
    """592 B per reverse step that also prepares the previous step's payload,
    512 B at a segment's first step; 200 steps in segments of 20 average 588."""
    assert bench.adj_particle_bytes_per_particle(200, 20) == pytest.approx((19 * 592 + 512) / 20)
    assert bench.adj_particle_bytes_per_particle(20, 20) == pytest.approx(588.0)
    # one segment per step: every step is a segment's first
    assert bench.adj_particle_bytes_per_particle(10, 1) == pytest.approx(512.0)
    # a ragged last segment counts its own steps
    assert bench.adj_particle_bytes_per_particle(25, 20) == pytest.approx(
        (19 * 592 + 512 + 4 * 592 + 512) / 25)


def test_g2p_byte_model_over_an_iteration():
    """G2P launches over one iteration: the forward pass (kept last segment
    writes records, others x and F only, each segment's last step the full
    state) and the recompute of every other segment (records; its last step
    runs no G2P)."""
    base, nxt = 48, 36 + 68
    full, record, xf = 96, 84, 48
    n_steps, K = 40, 20                      # two segments
    fwd = [xf + nxt] * 19 + [full] + [record + nxt] * 19 + [full]
    rec = [record + nxt] * 19
    expect = sum(base + b for b in fwd + rec) / (len(fwd) + len(rec))
    assert bench.g2p_bytes_per_particle(n_steps, K) == pytest.approx(expect)
