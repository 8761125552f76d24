"""State and config I/O in the reference's formats (python/morphompm/io.py:24-112).

    save_state(state, path)  .npz (lossless), .csv (lossless to 9 digits), .vtk (ParaView)
    load_state(path)         .npz, .csv
    save_config(cfg, path) / load_config(path)   .json
    load_curve(path, xcol, ycol)                 two numeric columns of a CSV

States may hold numpy arrays or device tensors (``ParticleState`` of the GPU
path); device buffers are copied to the host once per field.  The file
layouts are the reference's, so files move freely between the two
implementations: npz keys x, v, C, F, Fg; CSV columns x y z vx vy vz C00..C22
F00..F22 Fg00..Fg22 (row-major, 9 significant digits); legacy-ASCII VTK
POLYDATA with one vertex per particle and point scalars detF and detFg.
Unknown extensions raise ``ValueError``.  ``main()`` is the round-trip gate.
"""
import csv
import json
import os

import numpy as np
import torch

from .config import Material, SimConfig
from .state import ParticleState

FIELDS = ("x", "v", "C", "F", "Fg")
_MAT = [f"{i}{j}" for i in range(3) for j in range(3)]
CSV_COLUMNS = (["x", "y", "z", "vx", "vy", "vz"] + ["C" + s for s in _MAT] +
               ["F" + s for s in _MAT] + ["Fg" + s for s in _MAT])


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().to("cpu", torch.float64).numpy()
    return np.asarray(a, dtype=np.float64)


def _fields(state):
    return {f: _host(getattr(state, f)) for f in FIELDS}


def _ext(path):
    return os.path.splitext(path)[1].lower()


def save_state(state, path):
    ext = _ext(path)
    if ext == ".npz":
        np.savez(path, **_fields(state))
    elif ext == ".csv":
        _write_csv(_fields(state), path)
    elif ext == ".vtk":
        _write_vtk(_fields(state), path)
    else:
        raise ValueError(f"unsupported save format '{ext}' (use .npz/.csv/.vtk)")


def load_state(path):
    ext = _ext(path)
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as d:
            return ParticleState(*(np.array(d[f]) for f in FIELDS))
    if ext == ".csv":
        return _read_csv(path)
    raise ValueError(f"unsupported load format '{ext}' (use .npz/.csv)")


def _write_csv(f, path):
    n = f["x"].shape[0]
    table = np.concatenate([f["x"], f["v"], f["C"].reshape(n, 9), f["F"].reshape(n, 9),
                            f["Fg"].reshape(n, 9)], axis=1)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(CSV_COLUMNS)
        w.writerows([[f"{v:.9g}" for v in row] for row in table])


def _read_csv(path):
    with open(path, newline="") as fh:
        rows = list(csv.DictReader(fh))
    n = len(rows)
    table = np.array([[float(r[c]) for c in CSV_COLUMNS] for r in rows]).reshape(n, 33)
    return ParticleState(table[:, 0:3].copy(), table[:, 3:6].copy(),
                         table[:, 6:15].reshape(n, 3, 3), table[:, 15:24].reshape(n, 3, 3),
                         table[:, 24:33].reshape(n, 3, 3))


def _write_vtk(f, path):
    n = f["x"].shape[0]
    out = ["# vtk DataFile Version 3.0", "morphompm particles", "ASCII", "DATASET POLYDATA",
           f"POINTS {n} float"]
    out += [f"{p[0]:.6g} {p[1]:.6g} {p[2]:.6g}" for p in f["x"]]
    out.append(f"VERTICES {n} {2 * n}")
    out += [f"1 {p}" for p in range(n)]
    out.append(f"POINT_DATA {n}")
    for name in ("F", "Fg"):
        out += [f"SCALARS det{name} float 1", "LOOKUP_TABLE default"]
        out += [f"{v:.6g}" for v in np.linalg.det(f[name])]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def save_config(cfg, path):
    m = cfg.material
    d = {"N": cfg.N, "dx": cfg.dx, "dt": cfg.dt, "damping": cfg.damping, "sp": cfg.sp,
         "material": {"E": m.E, "nu": m.nu, "density": m.density}}
    if tuple(cfg.shape) != (cfg.N,) * 3:
        d["dims"] = list(cfg.shape)            # non-cubic grids (an extension)
    with open(path, "w") as fh:
        json.dump(d, fh, indent=2)


def load_config(path):
    with open(path) as fh:
        d = json.load(fh)
    m = d.get("material", {})
    mat = Material(E=m.get("E", 1e4), nu=m.get("nu", 0.3), density=m.get("density", 1000.0))
    kw = {"dims": tuple(d["dims"])} if "dims" in d else {}
    return SimConfig(N=d["N"], dx=d["dx"], dt=d["dt"], damping=d["damping"], sp=d["sp"],
                     material=mat, **kw)


def load_curve(path, xcol=0, ycol=1):
    """Rows whose two columns parse as numbers (headers and blanks skipped)."""
    xs, ys = [], []
    with open(path, newline="") as fh:
        for r in csv.reader(fh):
            try:
                x, y = float(r[xcol]), float(r[ycol])
            except (ValueError, IndexError):
                continue
            xs.append(x)
            ys.append(y)
    return np.array(xs), np.array(ys)


def main(device=None):
    """Round-trip gate (python/morphompm/io.py:116-148): npz exact, csv to
    1e-6, json config, vtk written.  With ``device`` the state lives in device
    memory.  Returns the number of failed checks."""
    import tempfile
    rng = np.random.default_rng(0)
    n = 20
    arrs = (rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
            rng.standard_normal((n, 3, 3)),
            np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.standard_normal((n, 3, 3)),
            np.tile(1.2 * np.eye(3), (n, 1, 1)))
    st = ParticleState(*(torch.tensor(a, device=device) if device else a for a in arrs))
    ref = dict(zip(FIELDS, arrs))
    fails = 0
    with tempfile.TemporaryDirectory() as d:
        for ext, atol in ((".npz", 0.0), (".csv", 1e-6)):
            p = os.path.join(d, "s" + ext)
            save_state(st, p)
            back = load_state(p)
            ok = all(np.allclose(ref[f], getattr(back, f), rtol=0, atol=atol) for f in FIELDS)
            print(f"    {ext}: x/v/C/F/Fg round trip [{'ok' if ok else 'FAIL'}]")
            fails += not ok
        cfg = SimConfig(N=12, dx=0.05)
        save_config(cfg, os.path.join(d, "c.json"))
        c2 = load_config(os.path.join(d, "c.json"))
        okc = c2.N == cfg.N and abs(c2.material.mu - cfg.material.mu) < 1e-9
        print(f"    .json: config round trip [{'ok' if okc else 'FAIL'}]")
        fails += not okc
        pv = os.path.join(d, "s.vtk")
        save_state(st, pv)
        okv = os.path.getsize(pv) > 0
        print(f"    .vtk: written [{'ok' if okv else 'FAIL'}]")
        fails += not okv
    return fails
