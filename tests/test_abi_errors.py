"""CPU tests of the error paths: every C ABI entry point rejects a bad batch
description or bad buffers with a negative MPM_E* code and a message, before
any device work, and the Python plan rejects instances the kernels cannot run.
Descriptors here carry placeholder table pointers that the library never
dereferences: every case below returns from host-side validation."""
import ctypes

import pytest

from morphompm_amd import _lib
from morphompm_amd._lib import MpmBatch, MpmError

MPM_EINVAL, MPM_ENOTSUP = -1, -3
FAKE = 0x1000          # non-NULL placeholder, never dereferenced


def _batch(**kw):
    """A descriptor that passes validation: 2 instances, tile 4x8x8."""
    b = MpmBatch()
    b.n_inst = 2
    b.tile[0], b.tile[1], b.tile[2] = 4, 8, 8
    b.n_slots, b.n_nodes, b.n_tiles = 512, 1024, 6
    b.n_bins = b.n_tiles * 4 * 8 * 8
    b.block = 256
    b.model_mask = 1
    b.inst = b.slot_inst = b.node_inst = b.tile_inst = FAKE
    for k, v in kw.items():
        if k == "tile":
            b.tile[0], b.tile[1], b.tile[2] = v
        elif k == "model_chunk_off":
            for m in range(4):
                b.model_chunk_off[m] = v[m]
        else:
            setattr(b, k, v)
    return b


def _ws(b, dtype=0):
    return int(_lib.lib().mpm_workspace_bytes(ctypes.byref(b), dtype))


def _msg():
    return _lib.lib().mpm_last_error().decode()


def test_valid_descriptor_sizes_workspace():
    b = _batch()
    n32, n64 = _ws(b, 0), _ws(b, 1)
    assert n32 > 0 and n64 > n32 and n32 % 256 == 0
    b2 = _batch(n_slots=1024)
    assert _ws(b2, 0) > n32          # more slots, more workspace


@pytest.mark.parametrize("kw, code, text", [
    (dict(n_inst=0), MPM_EINVAL, "batch tables missing"),
    (dict(inst=None), MPM_EINVAL, "batch tables missing"),
    (dict(tile_inst=None), MPM_EINVAL, "batch tables missing"),
    (dict(n_slots=300), MPM_EINVAL, "positive multiples of 256"),
    (dict(n_slots=0), MPM_EINVAL, "positive multiples of 256"),
    (dict(n_nodes=100), MPM_EINVAL, "positive multiples of 256"),
    (dict(n_slots=1 << 31), MPM_EINVAL, "below 2^31"),
    (dict(n_tiles=1), MPM_EINVAL, "dead unit"),
    (dict(n_bins=6 * 4 * 8 * 8 + 1), MPM_EINVAL, "n_tiles * tile cells"),
    (dict(block=100), MPM_EINVAL, "block must be 64, 128 or 256"),
    (dict(tile=(3, 3, 3), n_bins=6 * 27), MPM_ENOTSUP, "not built"),
    (dict(model_chunk=FAKE, model_chunk_off=(0, 1, 1, 3)), MPM_EINVAL, "model_chunk_off"),
    (dict(model_chunk=FAKE, model_chunk_off=(1, 1, 2, 2)), MPM_EINVAL, "model_chunk_off"),
    (dict(model_chunk=FAKE, model_chunk_off=(0, 2, 1, 2)), MPM_EINVAL, "model_chunk_off"),
])
def test_invalid_descriptor_rejected(kw, code, text):
    b = _batch(**kw)
    assert _ws(b) == code
    assert text in _msg()


def test_valid_model_chunks_accepted():
    assert _ws(_batch(model_chunk=FAKE, model_chunk_off=(0, 1, 1, 2))) > 0


def test_null_batch_and_bad_dtype():
    L = _lib.lib()
    assert L.mpm_workspace_bytes(None, 0) == MPM_EINVAL
    assert "NULL" in _msg()
    assert _ws(_batch(), 2) == MPM_EINVAL
    assert "dtype" in _msg()


def _fwd_args(b, **bufs):
    names = ("dyn_in", "Fg", "sched", "dyn_out", "grid", "sortbuf", "payload")
    vals = {n: FAKE + 256 * k for k, n in enumerate(names)}
    vals["sched"] = None
    vals.update(bufs)
    return [ctypes.byref(b)] + [vals[n] for n in names] + [1, bufs.get("ws", FAKE + 4096), None]


@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_step_fwd_rejects_bad_buffers(suf):
    L = _lib.lib()
    f = getattr(L, f"mpm_step_fwd_{suf}")
    assert f(*_fwd_args(_batch(block=7))) == MPM_EINVAL          # descriptor checked first
    assert "block" in _msg()
    assert f(*_fwd_args(_batch(), grid=None)) == MPM_EINVAL
    assert "NULL buffer" in _msg()
    assert f(*_fwd_args(_batch(), ws=None)) == MPM_EINVAL
    assert f(*_fwd_args(_batch(), dyn_out=FAKE)) == MPM_EINVAL   # aliases dyn_in
    assert "alias" in _msg()
    g = getattr(L, f"mpm_step_fwd_chain_{suf}")
    b = _batch()
    args = [FAKE, FAKE + 256, None, None, FAKE + 512, FAKE + 768, FAKE + 1024, FAKE + 1280,
            FAKE + 1280, 1, 0, 0, FAKE + 4096, None]                    # payload_next == payload
    assert g(ctypes.byref(b), *args) == MPM_EINVAL and "payload_next" in _msg()
    args[8] = FAKE + 1536
    for out in (-1, 4):                                               # MPM_OUT_* only
        args[11] = out
        assert g(ctypes.byref(b), *args) == MPM_EINVAL and "MPM_OUT" in _msg()
    args[11] = 3                                      # MPM_OUT_NONE ends a chain
    assert g(ctypes.byref(b), *args) == MPM_EINVAL and "payload_next must be NULL" in _msg()
    args[4] = None                                    # without G2P, dyn_out may be NULL ...
    args[8] = None
    args[11] = 2                                      # ... but not otherwise
    assert g(ctypes.byref(b), *args) == MPM_EINVAL and "NULL buffer" in _msg()


@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_step_bwd_rejects_bad_buffers(suf):
    L = _lib.lib()
    f = getattr(L, f"mpm_step_bwd_{suf}")
    b = _batch()
    base = [FAKE + 256 * k for k in range(11)]       # dyn_in .. bar_in, Fg_bar, theta_acc
    base[2] = None                                   # sched (optional)

    def call(fn, bufs, extra):
        return fn(ctypes.byref(b), *bufs, *extra)
    bufs = list(base)
    bufs[8] = bufs[7]                                # bar_in aliases bar_out
    assert call(f, bufs, [1, FAKE + 8192, None]) == MPM_EINVAL and "alias" in _msg()
    bufs = list(base)
    bufs[9] = None                                   # Fg_bar
    assert call(f, bufs, [1, FAKE + 8192, None]) == MPM_EINVAL and "NULL buffer" in _msg()
    g = getattr(L, f"mpm_step_bwd_chain_{suf}")
    # dyn_prev without sortbuf_prev
    assert call(g, base, [1, 0, FAKE + 9000, None, FAKE + 8192, None]) == MPM_EINVAL
    assert "prev" in _msg()
    # the payload area of a chained call is 0 (computed), 1 or 2
    assert call(g, base, [1, 3, None, None, FAKE + 8192, None]) == MPM_EINVAL
    assert "ready" in _msg()


@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_stress_entry_points_reject_bad_arguments(suf):
    L = _lib.lib()
    prm = (ctypes.c_double * 6)(1, 1, 0, 0, 0, 0)
    st = getattr(L, f"mpm_stress_{suf}")
    vj = getattr(L, f"mpm_stress_vjp_{suf}")
    assert st(7, prm, 4, 4, FAKE, FAKE, None, FAKE, None) == MPM_EINVAL
    assert "unknown model" in _msg()
    assert st(0, prm, -1, 4, FAKE, FAKE, None, FAKE, None) == MPM_EINVAL
    assert st(0, prm, 8, 4, FAKE, FAKE, None, FAKE, None) == MPM_EINVAL   # stride < n
    assert st(2, prm, 4, 4, FAKE, FAKE, None, FAKE, None) == MPM_EINVAL
    assert "HB needs C" in _msg()
    assert st(0, None, 4, 4, FAKE, FAKE, None, FAKE, None) == MPM_EINVAL
    assert st(0, prm, 0, 0, FAKE, FAKE, None, FAKE, None) == 0          # empty: nothing to do
    assert vj(1, prm, 4, 4, FAKE, FAKE, None, FAKE, None, FAKE, None, None, None) == MPM_EINVAL
    assert vj(2, prm, 4, 4, FAKE, FAKE, None, FAKE, FAKE, FAKE, None, None, None) == MPM_EINVAL
    assert "HB needs C" in _msg()
    assert vj(0, prm, 0, 0, FAKE, FAKE, None, FAKE, FAKE, FAKE, None, None, None) == 0


@pytest.mark.parametrize("suf", ["f32", "f64"])
def test_inversion_entry_points_reject_bad_arguments(suf):
    L = _lib.lib()
    b = _batch()
    f = getattr(L, f"mpm_growth_from_rates_{suf}")
    assert f(ctypes.byref(_batch(n_slots=300)), FAKE, FAKE, FAKE, None) == MPM_EINVAL
    assert "multiples" in _msg()
    assert f(ctypes.byref(b), None, FAKE, FAKE, None) == MPM_EINVAL and "NULL" in _msg()
    g = getattr(L, f"mpm_loss_det_{suf}")
    assert g(ctypes.byref(b), FAKE, FAKE + 256, None, FAKE, None) == MPM_EINVAL
    assert "NULL" in _msg()
    assert g(ctypes.byref(b), FAKE, FAKE, FAKE + 512, FAKE + 768, None) == MPM_EINVAL
    assert "alias" in _msg()
    h = getattr(L, f"mpm_growth_grad_{suf}")
    assert h(ctypes.byref(b), FAKE, None, FAKE, FAKE, None) == MPM_EINVAL and "NULL" in _msg()


def test_noalloc_step_rejects_batches_beyond_its_scratch():
    """The allocation-free forward step refuses, before any device call, a
    batch whose forward workspace exceeds the library's 64 MiB scratch."""
    L = _lib.lib()
    big = _batch(n_slots=1 << 24)                  # 2 x 64 MiB of sort keys alone
    args = [FAKE, FAKE + 256, None, FAKE + 512, FAKE + 768, FAKE + 1024, FAKE + 1280, 1, None]
    assert L.mpm_step_fwd_noalloc_f32(ctypes.byref(big), *args) == MPM_EINVAL
    assert "scratch" in _msg() and str(64 << 20) in _msg()
    assert L.mpm_step_fwd_noalloc_f32(ctypes.byref(_batch(block=7)), *args) == MPM_EINVAL
    assert "block" in _msg()


def test_gauss_newton_rejects_bad_arguments():
    L = _lib.lib()
    args = [4, FAKE, FAKE, FAKE, 1e-9, FAKE, FAKE, FAKE, None, 0, None, None]
    a = list(args)
    a[0] = 0
    assert L.mpm_gauss_newton(*a) == MPM_EINVAL and "n_inst" in _msg()
    a = list(args)
    a[5] = None
    assert L.mpm_gauss_newton(*a) == MPM_EINVAL and "NULL" in _msg()
    a = list(args)
    a[8] = FAKE                      # a history with no capacity or length counters
    assert L.mpm_gauss_newton(*a) == MPM_EINVAL and "hist_cap" in _msg()
    a = list(args)
    a[4] = -1.0
    assert L.mpm_gauss_newton(*a) == MPM_EINVAL and "tol" in _msg()


def test_kernel_timer_arguments():
    L = _lib.lib()
    assert L.mpm_kernel_timer(0, None, 0) == MPM_EINVAL and "events" in _msg()
    assert L.mpm_kernel_timer(-1, None, 0) == 0          # timer off


def test_check_raises_with_library_message():
    _ws(_batch(block=3))
    with pytest.raises(MpmError, match="block must be 64, 128 or 256"):
        _lib.check(MPM_EINVAL, "workspace_bytes")


# --------------------------------------------------------- Python-side checks
def _spec(**kw):
    from morphompm_amd import constitutive as K
    from morphompm_amd.config import Material, SimConfig
    from morphompm_amd.engine import InstanceSpec
    cfg = SimConfig(**{k: v for k, v in kw.items() if k in ("N", "dims", "dx", "dt")})
    s = InstanceSpec.from_cfg(kw.get("n", 10), cfg, K.NeoHookean(Material()))
    for k in ("pmass", "vol0"):
        if k in kw:
            setattr(s, k, kw[k])
    return s


@pytest.mark.parametrize("kw, text", [
    (dict(dims=(1023, 8, 8)), "1..1022"),
    (dict(dims=(8, 0, 8)), "1..1022"),
    (dict(n=-1), "negative"),
    (dict(dx=0.0), "dx"),
    (dict(dt=float("nan")), "dt"),
    (dict(pmass=0.0), "pmass"),
    (dict(vol0=-1.0), "vol0"),
])
def test_plan_rejects_unrunnable_instances(kw, text):
    from morphompm_amd.engine import BatchPlan
    with pytest.raises(ValueError, match=text):
        BatchPlan([_spec(N=8), _spec(**kw)])


def test_plan_rejects_empty_batch():
    from morphompm_amd.engine import BatchPlan
    with pytest.raises(ValueError, match="at least one instance"):
        BatchPlan([])


def test_largest_grid_passes_spec_check():
    from morphompm_amd.engine import MAX_AXIS_NODES, check_spec
    check_spec(_spec(dims=(MAX_AXIS_NODES, 1, MAX_AXIS_NODES)))
    assert MAX_AXIS_NODES + 1 == 1023          # base + 2 of the last bin fits 10 bits


def test_model_without_device_implementation_raises():
    from morphompm_amd.engine import model_spec

    class Plastic:
        kind = "plastic"
    with pytest.raises(NotImplementedError):
        model_spec(Plastic())
