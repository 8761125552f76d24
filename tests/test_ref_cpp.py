"""The reference's own C++ forward solver (oracle/_ref, built from the
reference sources by oracle/Makefile) as a second forward checker, with the
tolerances of the reference's numpy <-> C++ parity gate (scripts/parity.py:48:
max|dF| < 1e-2, relative mean det F < 1e-3).  The C++ solver clamps singular
values at 1e-6, masks grid mass at 1e-9 and skips zero-weight nodes; these
scenes never reach those branches."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import ref_loop as RL
from oracle import ref_vec as RV

LIB = os.path.join(ROOT, "oracle", "_ref", "libref_growth.so")
_P = ctypes.c_void_p
need_ref = pytest.mark.skipif(not os.path.exists(LIB),
                              reason="oracle/_ref not built (needs the reference sources)")


def _lib():
    L = ctypes.CDLL(LIB)
    L.ref_kirchhoff.argtypes = [_P, _P, ctypes.c_float, ctypes.c_float, _P, _P]
    L.ref_run.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P, _P,
                          ctypes.c_int, ctypes.c_float, _P, _P]
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def cpp_run(pts, Fg, N, dx, steps, dt, E=1e4, nu=0.3, damping=0.05, density=1000.0, sp=0.03):
    n = len(pts)
    xo = np.zeros((n, 3), np.float32)
    Fo = np.zeros((n, 3, 3), np.float32)
    x, G = _f32(pts), _f32(Fg)
    _lib().ref_run(N, dx, E, nu, damping, density, sp, n, x.ctypes.data_as(_P),
                   G.ctypes.data_as(_P), steps, dt, xo.ctypes.data_as(_P), Fo.ctypes.data_as(_P))
    return xo.astype(np.float64), Fo.astype(np.float64)


def parity_scene():
    """tests/test_growth.cpp:337-357: 27-particle Hencky free swell, Fg = 1.05 I,
    N = 12, dx = 0.05, 30 steps of 1e-3."""
    c = np.arange(0.27, 0.33 + 1e-9, 0.03)
    pts = np.array([(x, y, z) for x in c for y in c for z in c])
    Fg = np.repeat((1.05 * np.eye(3))[None], len(pts), 0)
    cfg = RL.SimConfig(N=12, dx=0.05, dt=1.0e-3, damping=0.05, sp=0.03)
    return pts, Fg, cfg


@need_ref
def test_hencky_stress_matches_cpp():
    mat = RL.Material()
    hk = RL.Hencky(mat)
    rng = np.random.default_rng(3)
    L = _lib()
    for _ in range(20):
        F = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
        Fg = np.diag(rng.uniform(0.8, 1.3, 3))
        tau = np.zeros(9, np.float32)
        a, b = _f32(F), _f32(Fg)
        L.ref_kirchhoff(a.ctypes.data_as(_P), b.ctypes.data_as(_P), mat.E, mat.nu,
                        tau.ctypes.data_as(_P), None)
        ref = hk.stress(F, Fg)
        assert np.abs(tau.reshape(3, 3) - ref).max() <= 2e-4 * np.abs(ref).max() + 1e-2


@need_ref
def test_oracle_parity_scene_matches_cpp():
    pts, Fg, cfg = parity_scene()
    st = RL.ParticleState.rest(pts)
    st.Fg = Fg.copy()
    fin, _ = RV.rollout(st, RL.Hencky(cfg.material), cfg, 30, advect=True, keep=False)
    xc, Fc = cpp_run(pts, Fg, 12, 0.05, 30, 1e-3)
    assert np.abs(fin.F - Fc).max() < 1e-2
    dn, dc = np.linalg.det(fin.F).mean(), np.linalg.det(Fc).mean()
    assert abs(dn - dc) / abs(dc) < 1e-3


CPP_FIXTURE = os.path.join(ROOT, "tests", "golden", "cpp_parity_scene.npz")


@need_ref
def test_cpp_fixture_is_the_solver_output():
    """tests/golden/cpp_parity_scene.npz (written by make_cpp_fixture.py) holds
    exactly what the reference's C++ solver computes for its parity scene."""
    pts, Fg, _ = parity_scene()
    with np.load(CPP_FIXTURE) as d:
        assert np.array_equal(d["pts"], pts) and np.array_equal(d["Fg"], Fg)
        xc, Fc = cpp_run(pts, Fg, 12, 0.05, int(d["steps"]), 1e-3)
        assert np.array_equal(d["x"], xc) and np.array_equal(d["F"], Fc)


@pytest.mark.gpu
def test_gpu_fp32_parity_scene_matches_cpp():
    """The GPU FP32 trajectory of the same scene against the C++ FP32 solver's
    output (the committed fixture: no compiled reference code on the GPU side)."""
    from morphompm_amd import constitutive as K, integrate as G
    from morphompm_amd.config import SimConfig
    from morphompm_amd.state import ParticleState
    with np.load(CPP_FIXTURE) as d:
        pts, Fg, steps, xc, Fc = d["pts"], d["Fg"], int(d["steps"]), d["x"], d["F"]
    cfg = SimConfig(N=12, dx=0.05, dt=1.0e-3, damping=0.05, sp=0.03)
    st = ParticleState.rest(pts.astype(np.float32))
    st.Fg = Fg.astype(np.float32)
    fin, _ = G.rollout(st, K.Hencky(cfg.material), cfg, steps, advect=True)
    assert np.abs(fin.F - Fc).max() < 1e-2
    assert abs(np.linalg.det(fin.F).mean() - np.linalg.det(Fc).mean()) / abs(np.linalg.det(Fc).mean()) < 1e-3
    assert np.abs(fin.x - xc).max() < 1e-3


@need_ref
def test_timoshenko_scene_reproduces_fixture():
    """The reference's bilayer-bending scene run by its own C++ solver gives the
    curvatures of outputs/figures/bending_timoshenko.csv (copied to
    tests/golden): this pins the scene and the curvature fit the GPU test uses."""
    from morphompm_amd.growth import bilayer_curvature
    from morphompm_amd.scenes import timoshenko_bilayer
    fx = np.loadtxt(os.path.join(ROOT, "tests", "golden", "bending_timoshenko.csv"),
                    delimiter=",", skiprows=1)
    for eps, k_sim, k_t in fx:
        sc = timoshenko_bilayer(eps)
        x, _ = cpp_run(sc.pts, sc.G, 48, 0.025, 4000, 1e-3, E=1e4, nu=0.3, damping=0.05,
                       density=1000.0, sp=0.0125)
        assert abs(bilayer_curvature(x) - k_sim) <= 1e-3
        assert sc.extra["kappa_timoshenko"] == pytest.approx(k_t)
