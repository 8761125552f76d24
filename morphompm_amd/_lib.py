"""ctypes binding of libmorphompm_hip.so (the C ABI in include/morphompm_hip.h).

The HIP library is the only compute path: if it is missing or cannot be
loaded, every entry point raises.  There is no CPU fallback.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
# MORPHOMPM_AMD_LIB names another build of the same library (same ABI), e.g.
# to compare two builds on one device; the in-tree build is the default.
LIB_PATH = os.environ.get("MORPHOMPM_AMD_LIB") or os.path.join(HERE, "_lib", "libmorphompm_hip.so")

ABI_VERSION = 7
SLOT_ALIGN = 256

NEO_HOOKEAN, HENCKY, HERSCHEL_BULKLEY = 0, 1, 2
# rows of dyn_out a chained step writes; OUT_NONE: no G2P (the step record only)
OUT_FULL, OUT_RECORD, OUT_NEXT, OUT_NONE = 0, 1, 2, 3

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


class MpmInstance(ctypes.Structure):
    _fields_ = [("p_off", _I64), ("g_off", _I64), ("reserved", _I64), ("t_off", _I64),
                ("n", _I32), ("nx", _I32), ("ny", _I32), ("nz", _I32),
                ("tx", _I32), ("ty", _I32), ("tz", _I32), ("model", _I32),
                ("shift", _I32 * 3), ("pad", _I32), ("origin", _I32 * 3), ("pad2", _I32),
                ("dx", ctypes.c_double), ("dt", ctypes.c_double),
                ("damping", ctypes.c_double), ("pmass", ctypes.c_double),
                ("vol0", ctypes.c_double), ("prm", ctypes.c_double * 6)]


class MpmBatch(ctypes.Structure):
    _fields_ = [("n_inst", _I32), ("tile", _I32 * 3), ("n_slots", _I64), ("n_nodes", _I64),
                ("n_bins", _I64), ("n_tiles", _I64), ("block", _I32), ("model_mask", _I32),
                ("inst", _P), ("slot_inst", _P), ("node_inst", _P), ("tile_inst", _P),
                ("model_chunk", _P), ("model_chunk_off", _I64 * 4)]


class MpmError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library; raises if it is not built or its ABI differs."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MpmError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.mpm_abi_version.restype = _I32
    L.mpm_last_error.restype = ctypes.c_char_p
    L.mpm_workspace_bytes.restype = _I64
    L.mpm_workspace_bytes.argtypes = [ctypes.POINTER(MpmBatch), _I32]
    L.mpm_tile_shapes.restype = _I32
    L.mpm_tile_shapes.argtypes = [ctypes.POINTER(_I32), _I32]
    for suf in ("f32", "f64"):
        f = getattr(L, f"mpm_stress_{suf}")
        f.restype = _I32
        f.argtypes = [_I32, ctypes.POINTER(ctypes.c_double), _I64, _I64, _P, _P, _P, _P, _P]
        f = getattr(L, f"mpm_stress_vjp_{suf}")
        f.restype = _I32
        f.argtypes = [_I32, ctypes.POINTER(ctypes.c_double), _I64, _I64] + [_P] * 9
        f = getattr(L, f"mpm_sort_by_cell_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P]
        f = getattr(L, f"mpm_step_fwd_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P]
        f = getattr(L, f"mpm_step_fwd_chain_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch)] + [_P] * 9 + [_I32, _I32, _I32, _P, _P]
        f = getattr(L, f"mpm_step_bwd_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch)] + [_P] * 11 + [_I32, _P, _P]
        f = getattr(L, f"mpm_step_bwd_chain_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch)] + [_P] * 11 + [_I32, _I32, _P, _P, _P, _P]
        f = getattr(L, f"mpm_growth_from_rates_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P]
        f = getattr(L, f"mpm_loss_det_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P, _P]
        f = getattr(L, f"mpm_growth_grad_{suf}")
        f.restype = _I32
        f.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P, _P]
    L.mpm_step_fwd_noalloc_f32.restype = _I32
    L.mpm_step_fwd_noalloc_f32.argtypes = [ctypes.POINTER(MpmBatch), _P, _P, _P, _P, _P, _P, _P,
                                           _I32, _P]
    L.mpm_gauss_newton.restype = _I32
    L.mpm_gauss_newton.argtypes = [_I32, _P, _P, _P, ctypes.c_double, _P, _P, _P, _P, _I32, _P,
                                   _P]
    L.mpm_kernel_timer.restype = _I32
    L.mpm_kernel_timer.argtypes = [_I32, _P, _I32]
    L.mpm_kernel_timer_count.restype = _I32
    if L.mpm_abi_version() != ABI_VERSION:
        raise MpmError(f"ABI mismatch: library {L.mpm_abi_version()} vs binding {ABI_VERSION}")
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = lib().mpm_last_error().decode(errors="replace")
        raise MpmError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def suffix(dtype):
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise MpmError(f"unsupported dtype {dtype} (float32 or float64)")


def tile_shapes():
    L = lib()
    n = L.mpm_tile_shapes(None, 0)
    buf = (_I32 * (3 * n))()
    L.mpm_tile_shapes(buf, n)
    return [tuple(buf[3 * i:3 * i + 3]) for i in range(n)]


def require_gpu():
    if not torch.cuda.is_available():
        raise MpmError("no HIP device visible: the MI355X path needs a GPU")
    lib()
