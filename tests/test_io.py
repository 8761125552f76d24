"""State / config I/O in the reference's formats (python/morphompm/io.py)."""
import csv
import os

import numpy as np
import pytest

from morphompm_amd import io
from morphompm_amd.config import SimConfig
from morphompm_amd.state import ParticleState


def _state(n=7, seed=1):
    rng = np.random.default_rng(seed)
    return ParticleState(rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                         rng.standard_normal((n, 3, 3)),
                         np.eye(3) + 0.1 * rng.standard_normal((n, 3, 3)),
                         np.tile(np.diag([1.1, 1.2, 0.9]), (n, 1, 1)))


def test_round_trip_gate():
    assert io.main() == 0


def test_csv_layout_is_the_references(tmp_path):
    """Header x y z vx vy vz C00..C22 F00..F22 Fg00..Fg22 (io.py:20-24), row-major."""
    st = _state()
    p = str(tmp_path / "s.csv")
    io.save_state(st, p)
    with open(p) as f:
        rows = list(csv.reader(f))
    hdr = rows[0]
    assert hdr[:6] == ["x", "y", "z", "vx", "vy", "vz"] and len(hdr) == 33
    assert hdr[6] == "C00" and hdr[7] == "C01" and hdr[14] == "C22"
    assert hdr[15] == "F00" and hdr[24] == "Fg00" and hdr[32] == "Fg22"
    assert float(rows[1][7]) == pytest.approx(st.C[0, 0, 1], rel=1e-8)
    assert float(rows[1][24 + 5]) == pytest.approx(st.Fg[0, 1, 2], rel=1e-8)


def test_npz_keys_and_exact(tmp_path):
    st = _state()
    p = str(tmp_path / "s.npz")
    io.save_state(st, p)
    with np.load(p) as d:
        assert sorted(d.files) == sorted(io.FIELDS)
    back = io.load_state(p)
    for f in io.FIELDS:
        assert np.array_equal(getattr(back, f), getattr(st, f))


def test_vtk_points_and_dets(tmp_path):
    st = _state(5)
    p = str(tmp_path / "s.vtk")
    io.save_state(st, p)
    lines = open(p).read().splitlines()
    assert lines[0].startswith("# vtk DataFile") and lines[3] == "DATASET POLYDATA"
    assert lines[4] == "POINTS 5 float" and lines[10] == "VERTICES 5 10"
    i = lines.index("SCALARS detF float 1")
    assert float(lines[i + 2]) == pytest.approx(np.linalg.det(st.F[0]), rel=1e-5)


def test_config_and_curve(tmp_path):
    cfg = SimConfig(N=16, dx=1 / 16, dt=2e-4, damping=0.1, sp=1 / 32, dims=(16, 8, 8))
    p = str(tmp_path / "c.json")
    io.save_config(cfg, p)
    c2 = io.load_config(p)
    assert c2.shape == (16, 8, 8) and c2.dt == cfg.dt and c2.material.E == cfg.material.E
    q = str(tmp_path / "k.csv")
    with open(q, "w") as f:
        f.write("eps,kappa\n0.02,0.359\n\n0.04,0.669\n")
    x, y = io.load_curve(q)
    assert np.array_equal(x, [0.02, 0.04]) and np.array_equal(y, [0.359, 0.669])


def test_unknown_extension_raises(tmp_path):
    with pytest.raises(ValueError):
        io.save_state(_state(), str(tmp_path / "s.h5"))
    with pytest.raises(ValueError):
        io.load_state(str(tmp_path / "s.vtk"))
