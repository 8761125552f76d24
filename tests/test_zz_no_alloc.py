"""mpm_step_fwd_noalloc_f32, the forward step that allocates nothing (its
scratch is an array the library holds on each GPU), against mpm_step_fwd_f32
with a caller workspace: bit for bit, on two calls in a row, the second with a
smaller batch (another scratch layout over the same array, with what the
first call left in it)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(cells, seed):
    from morphompm_amd import constitutive as K
    from morphompm_amd.engine import BatchPlan, InstanceSpec
    from morphompm_amd.scenes import free_swell_cube
    sc = free_swell_cube(cells=cells, seed=seed, model="hencky")
    plan = BatchPlan([InstanceSpec.from_cfg(sc.n, sc.cfg, K.Hencky(sc.cfg.material), sc.pts)])
    rng = np.random.default_rng(seed)
    n = sc.n
    F = np.eye(3) + 0.02 * rng.standard_normal((n, 3, 3))
    dyn = plan.pack_dyn([(sc.pts, 0.05 * rng.standard_normal((n, 3)),
                          0.5 * rng.standard_normal((n, 3, 3)), F)], torch.float32)
    G = plan.pack_mat([sc.G], torch.float32)
    return plan, dyn, G


def _outputs(plan):
    return (plan.new_dyn(torch.float32), plan.new_grid(torch.float32), plan.new_sort(),
            plan.new_payload(torch.float32))


def test_noalloc_step_matches_workspace_step():
    from morphompm_amd import _lib
    from morphompm_amd._lib import ptr, stream_ptr
    L = _lib.lib()
    for cells, seed in ((8, 0), (5, 1)):       # 4096 particles, then 1000
        plan, dyn, G = _case(cells, seed)
        print(f"noalloc step: {plan.specs[0].n} particles", flush=True)
        ref, out = _outputs(plan), _outputs(plan)
        ws = plan.workspace(torch.float32)
        rc = L.mpm_step_fwd_f32(ctypes.byref(plan.c), ptr(dyn), ptr(G), None, ptr(ref[0]),
                                ptr(ref[1]), ptr(ref[2]), ptr(ref[3]), 1, ptr(ws), stream_ptr())
        assert rc == 0, L.mpm_last_error()
        rc = L.mpm_step_fwd_noalloc_f32(ctypes.byref(plan.c), ptr(dyn), ptr(G), None,
                                        ptr(out[0]), ptr(out[1]), ptr(out[2]), ptr(out[3]), 1,
                                        stream_ptr())
        assert rc == 0, L.mpm_last_error()
        torch.cuda.synchronize()
        for name, a, b in zip(("state", "grid", "sort record", "payload"), ref, out):
            assert torch.equal(a, b), (cells, name)
        assert torch.isfinite(out[0][:, plan.live]).all()
