"""CPU tests: the C ABI library loads and exports every declared symbol, the
ctypes structs match the C layout, and the device math (mpm_math.h, built for
the host) matches the oracle and finite differences.  No GPU calls."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import ref_loop as RL
from oracle import ref_vec as RV

HEADER = os.path.join(ROOT, "include", "morphompm_hip.h")
HIP_LIB = os.path.join(ROOT, "morphompm_amd", "_lib", "libmorphompm_hip.so")
HOST_LIB = os.path.join(ROOT, "morphompm_amd", "_lib", "libmpm_host.so")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int64_t|const char\*)\s+(mpm_\w+)\(", txt,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 14
    lib = ctypes.CDLL(HIP_LIB)
    for n in names:
        assert hasattr(lib, n), n
    lib.mpm_abi_version.restype = ctypes.c_int32
    version = int(re.search(r"#define MPM_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert lib.mpm_abi_version() == version


def test_ctypes_structs_match_c_layout(tmp_path):
    from morphompm_amd._lib import MpmBatch, MpmInstance
    src = tmp_path / "layout.c"
    fields_i = [f[0] for f in MpmInstance._fields_]
    fields_b = [f[0] for f in MpmBatch._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "morphompm_hip.h"',
             "int main(void){",
             'printf("%zu %zu\\n", sizeof(mpm_instance), sizeof(mpm_batch));']
    lines += [f'printf("%zu\\n", offsetof(mpm_instance, {f}));' for f in fields_i]
    lines += [f'printf("%zu\\n", offsetof(mpm_batch, {f}));' for f in fields_b]
    lines += ["return 0;}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == ctypes.sizeof(MpmInstance) and vals[1] == ctypes.sizeof(MpmBatch)
    offs = vals[2:]
    assert offs[:len(fields_i)] == [getattr(MpmInstance, f).offset for f in fields_i]
    assert offs[len(fields_i):] == [getattr(MpmBatch, f).offset for f in fields_b]


# ------------------------------------------------------- host-built device math
_P = ctypes.c_void_p


def _host():
    lib = ctypes.CDLL(HOST_LIB)
    lib.mpm_host_stress.argtypes = [ctypes.c_int, ctypes.c_int, _P, ctypes.c_long] + [_P] * 4
    lib.mpm_host_stress_vjp.argtypes = [ctypes.c_int, ctypes.c_int, _P, ctypes.c_long] + [_P] * 8
    lib.mpm_host_sym_eig.argtypes = [ctypes.c_int, ctypes.c_long, _P, _P, _P]
    return lib


def _p(a):
    return a.ctypes.data_as(_P)


def _models():
    mat = RL.Material()
    return {"nh": (0, [mat.mu, mat.lam, 0, 0, 0, 0], RL.NeoHookean(mat)),
            "hencky": (1, [mat.mu, mat.lam, 0, 0, 0, 0], RL.Hencky(mat)),
            "hb": (2, [1e4, 1.0, 0.5, 50.0, 100.0, 1e-4], RL.HerschelBulkley())}


def host_stress(fp32, kind, prm, F, Fg, C):
    n = len(F)
    out = np.zeros((n, 3, 3))
    prm = np.array(prm, float)
    _host().mpm_host_stress(fp32, kind, _p(prm), n, _p(np.ascontiguousarray(F)),
                            _p(np.ascontiguousarray(Fg)), _p(np.ascontiguousarray(C)), _p(out))
    return out


def host_vjp(fp32, kind, prm, F, Fg, C, tb):
    n = len(F)
    Fb, Gb, Cb, th = np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.zeros((n, 2))
    prm = np.array(prm, float)
    _host().mpm_host_stress_vjp(fp32, kind, _p(prm), n, *(_p(np.ascontiguousarray(a))
                                                          for a in (F, Fg, C, tb)),
                                _p(Fb), _p(Gb), _p(Cb), _p(th))
    return Fb, Gb, Cb, th


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_host_math_golden(golden, name):
    kind, prm, _ = _models()[name]
    F, Fg, C, tb = (golden[f"cons_{k}"] for k in ("F", "Fg", "C", "tb"))
    assert _rel(host_stress(0, kind, prm, F, Fg, C), golden[f"cons_{name}_tau"]) < 1e-13
    Fb, Gb, Cb, _ = host_vjp(0, kind, prm, F, Fg, C, tb)
    assert _rel(Fb, golden[f"cons_{name}_Fb"]) < 1e-13
    assert _rel(Gb, golden[f"cons_{name}_Fgb"]) < 1e-13
    if name == "hb":
        assert _rel(Cb, golden[f"cons_{name}_Cb"]) < 1e-13
    # FP32 build of the same functions
    assert _rel(host_stress(1, kind, prm, F, Fg, C), golden[f"cons_{name}_tau"]) < 2e-6
    Fb, Gb, _, _ = host_vjp(1, kind, prm, F, Fg, C, tb)
    assert _rel(Fb, golden[f"cons_{name}_Fb"]) < 2e-6
    assert _rel(Gb, golden[f"cons_{name}_Fgb"]) < 2e-6


def test_host_hencky_degenerate(golden):
    kind, prm, _ = _models()["hencky"]
    F, tb = golden["deg_F"], golden["deg_tb"]
    I = np.repeat(np.eye(3)[None], len(F), 0)
    for fp32, tol in ((0, 1e-9), (1, 2e-6)):
        Fb, _, _, _ = host_vjp(fp32, kind, prm, F, I, np.zeros_like(F), tb)
        for i in range(len(F)):
            assert _rel(Fb[i], golden["deg_Fb"][i]) < tol, (fp32, i)


def test_host_hb_theta_and_small_rate():
    """(tau_y, K) cotangents match the oracle; FP32 stays accurate as gdot -> eps
    (the cancellation-free series branch)."""
    kind, prm, m = _models()["hb"]
    rng = np.random.default_rng(4)
    F = np.eye(3) + 0.05 * rng.standard_normal((8, 3, 3))
    Fg = 1.1 * np.eye(3) + 0.02 * rng.standard_normal((8, 3, 3))
    tb = rng.standard_normal((8, 3, 3))
    for scale in (1e-1, 1e-3, 1e-5):
        C = scale * rng.standard_normal((8, 3, 3))
        _, _, Cb, th = host_vjp(0, kind, prm, F, Fg, C, tb)
        assert _rel(th, RV.param_vjp(m, C, tb)) < 1e-12
        _, _, Cref = RV.stress_vjp(m, F, Fg, C, tb)
        _, _, Cb32, th32 = host_vjp(1, kind, prm, F, Fg, C, tb)
        assert _rel(Cb32, Cref) < 5e-5, scale
        assert _rel(th32, RV.param_vjp(m, C, tb)) < 5e-5, scale


def test_host_jacobi_eigen():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((64, 3, 3))
    S = A @ np.swapaxes(A, 1, 2) + 0.1 * np.eye(3)
    S[:4] = np.diag([2.0, 2.0, 1.0]) + 1e-9 * rng.standard_normal((4, 3, 3))   # near-repeated
    S = 0.5 * (S + np.swapaxes(S, 1, 2))
    for fp32, tol in ((0, 1e-13), (1, 2e-6)):
        lam = np.zeros((64, 3)); V = np.zeros((64, 3, 3))
        _host().mpm_host_sym_eig(fp32, 64, _p(np.ascontiguousarray(S)), _p(lam), _p(V))
        rec = V @ (lam[:, :, None] * np.swapaxes(V, 1, 2))
        assert np.abs(rec - S).max() / np.abs(S).max() < tol
        assert np.abs(np.swapaxes(V, 1, 2) @ V - np.eye(3)).max() < 10 * tol


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_host_vjp_fd(name):
    """Host-built FP64 VJP vs central FD of the host-built FP64 stress
    (constitutive.py:183-219 gate, tolerance 1e-5)."""
    kind, prm, _ = _models()[name]
    rng = np.random.default_rng(9)
    states = [(np.diag([1.20, 1.00, 0.85]) + 0.03 * np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]]),
               np.diag([1.30, 1.10, 0.90])),
              (np.eye(3) + 0.05 * rng.standard_normal((3, 3)), np.diag([1.4, 1.0, 0.8]))]
    eps = 1e-6
    for F, Fg in states:
        for _ in range(3):
            C = 0.5 * rng.standard_normal((3, 3)) if name == "hb" else np.zeros((3, 3))
            tb = rng.standard_normal((3, 3)); tb = 0.5 * (tb + tb.T)
            dF, dG, dC = (rng.standard_normal((3, 3)) for _ in range(3))
            if name != "hb":
                dC = np.zeros((3, 3))
            Fb, Gb, Cb, _ = host_vjp(0, kind, prm, F[None], Fg[None], C[None], tb[None])
            Lp = np.sum(tb * host_stress(0, kind, prm, (F + eps * dF)[None], (Fg + eps * dG)[None],
                                         (C + eps * dC)[None])[0])
            Lm = np.sum(tb * host_stress(0, kind, prm, (F - eps * dF)[None], (Fg - eps * dG)[None],
                                         (C - eps * dC)[None])[0])
            fd = (Lp - Lm) / (2 * eps)
            an = np.sum(Fb[0] * dF) + np.sum(Gb[0] * dG) + np.sum(Cb[0] * dC)
            assert abs(an - fd) <= 1e-5 * (abs(fd) + 1e-12)


def test_bin_keys_numpy_and_torch_agree():
    """The slot order of the batched and single-scene paths sorts particles by
    bin key; the key is computed with numpy on the host or with torch on the
    device, and both give the same integers (off-grid positions clamp to the
    border bins)."""
    import torch
    from morphompm_amd.engine import bin_keys
    rng = np.random.default_rng(0)
    for shape, tile, shift in [((20, 30, 17), (4, 4, 8), (1, 3, 5)), ((8, 8, 8), (2, 4, 8), (0, 0, 0)),
                               ((33, 9, 70), (8, 8, 4), (7, 2, 3))]:
        dx = 0.05
        x = rng.uniform(-0.2, 1.2, (5000, 3)) * np.array(shape) * dx
        a = bin_keys(x, dx, shape, tile, shift)
        b = bin_keys(torch.tensor(x), dx, shape, tile, shift).numpy()
        assert np.array_equal(a, b)
        T = np.array(tile)
        nt = -(-(np.array(shape) + 2 + np.array(shift)) // T)
        assert a.min() >= 0 and a.max() < int(np.prod(nt)) * int(np.prod(T))


def test_fp32_hencky_log_series_matches_fp64():
    """The FP32 Hencky path through the matrix-log series (no eigen-
    decomposition while |Z|_F <= 0.2) and its reverse-mode VJP against the
    FP64 eigen path, for deformations from exactly isotropic (where an eigen
    solver meets repeated eigenvalues) to stretches beyond the series' range
    (the FP32 eigen path takes over): stress within 3e-6, F and Fg cotangents
    within 2e-6 of the largest FP64 entry."""
    rng = np.random.default_rng(0)
    prm = np.array([384.6, 576.9, 0, 0, 0, 0.0])
    for sd in (0.0, 1e-7, 1e-4, 1e-2, 0.05, 0.1, 0.25):
        n = 200
        R = np.linalg.qr(rng.standard_normal((n, 3, 3)))[0]
        st = np.exp(sd * rng.standard_normal((n, 3)))
        F = R @ (st[:, :, None] * np.linalg.qr(rng.standard_normal((n, 3, 3)))[0])
        if sd == 0.0:
            F = np.repeat(np.eye(3)[None], n, 0) * 1.03
        Fg = np.repeat(np.eye(3)[None], n, 0) * rng.uniform(0.9, 1.1, (n, 1, 1))
        C = np.zeros((n, 3, 3))
        tb = rng.standard_normal((n, 3, 3))
        tb = tb + tb.transpose(0, 2, 1)
        assert _rel(host_stress(1, 1, prm, F, Fg, C), host_stress(0, 1, prm, F, Fg, C)) < 3e-6
        a, b = host_vjp(1, 1, prm, F, Fg, C, tb), host_vjp(0, 1, prm, F, Fg, C, tb)
        assert _rel(a[0], b[0]) < 2e-6 and _rel(a[1], b[1]) < 2e-6, sd
