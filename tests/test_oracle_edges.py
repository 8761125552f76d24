"""Pins the oracle's edge behaviour on the CPU: stencil nodes off the grid are
skipped (python/morphompm/transfer.py:47-51), nodes with grid mass at or below
the mass threshold get no velocity (transfer.py:55-58), and a particle with no
node on the grid keeps x and F and gets v = C = 0.  Each case is reduced to a
closed form of a single stress-free particle (F = Fg = I, C = 0, so tau = 0
and the APIC terms vanish): every live node then carries velocity s v0, and

    v' = s v0 S,   C' = k4 s v0 (sum_i w_i (x_i - x_p))^T,   S = sum_i w_i

over the live nodes i.  The weights come from the quadratic B-spline written
as a function of the node distance r (in cells), N(r) = 3/4 - r^2 for |r| <
1/2 and (3/2 - |r|)^2 / 2 for 1/2 <= |r| < 3/2, not from the oracle's own
(1.5 - fx) form.  The adjoint at the edge is pinned by central differences."""
import numpy as np
import pytest

from oracle import ref_loop as RL
from oracle import ref_vec as RV

CFG = RL.SimConfig(N=8, dx=0.1, dt=1e-3, damping=0.05, sp=0.02)
I3 = np.eye(3)


def _N(r):
    r = abs(r)
    if r < 0.5:
        return 0.75 - r * r
    if r < 1.5:
        return 0.5 * (1.5 - r) ** 2
    return 0.0


def _axis_sums(X, n_nodes, live=lambda i: True):
    """(sum_i N(i - X), sum_i N(i - X)(i - X)) over on-grid nodes i of one axis."""
    S = M = 0.0
    for i in range(int(np.floor(X)) - 2, int(np.floor(X)) + 3):
        if 0 <= i < n_nodes and live(i):
            w = _N(i - X)
            S += w
            M += w * (i - X)
    return S, M


def _closed_form(x, v0, cfg):
    X = x / cfg.dx
    sums = [_axis_sums(X[a], cfg.shape[a]) for a in range(3)]
    S = np.prod([s for s, _ in sums])
    s = 1.0 - cfg.damping
    k4 = 4.0 / cfg.dx ** 2
    m = np.array([sums[c][1] * np.prod([sums[a][0] for a in range(3) if a != c]) * cfg.dx
                  for c in range(3)])
    return s * v0 * S, k4 * s * np.outer(v0, m)


def _single(x, v0, F=None):
    return RL.ParticleState(np.array([x], float), np.array([v0], float), np.zeros((1, 3, 3)),
                            (I3 if F is None else F)[None].copy(), I3[None].copy())


def _models():
    mat = RL.Material()
    return {"nh": RL.NeoHookean(mat), "hencky": RL.Hencky(mat), "hb": RL.HerschelBulkley()}


STRADDLE = [(0.03, 0.455, 0.74), (0.785, 0.012, 0.33), (-0.07, 0.81, 0.5), (0.42, 0.37, 0.842)]


@pytest.mark.parametrize("impl", [RL, RV])
@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
@pytest.mark.parametrize("x", STRADDLE)
def test_straddling_particle_closed_form(impl, name, x):
    v0 = np.array([0.3, -0.2, 0.5])
    ns, _ = impl.step(_single(x, v0), _models()[name], CFG, advect=True)
    vn, Cn = _closed_form(np.array(x), v0, CFG)
    S_full = 1.0
    assert np.allclose(ns.v[0], vn, rtol=1e-13, atol=1e-15)
    assert not np.allclose(ns.v[0], (1 - CFG.damping) * v0 * S_full)   # really truncated
    assert np.allclose(ns.C[0], Cn, rtol=1e-12, atol=1e-13)
    assert np.allclose(ns.F[0], I3 + CFG.dt * Cn, rtol=1e-14, atol=1e-15)
    assert np.allclose(ns.x[0], np.array(x) + CFG.dt * vn, rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("impl", [RL, RV])
@pytest.mark.parametrize("x", [(-0.5, 0.4, 0.4), (0.4, 0.9, 0.4), (0.4, 0.4, -0.16),
                               (1e3, 0.4, 0.4)])
def test_particle_off_the_grid_keeps_x_and_F(impl, x):
    rng = np.random.default_rng(5)
    F = I3 + 0.1 * rng.standard_normal((3, 3))
    st = _single(x, np.array([1.0, 2.0, 3.0]), F)
    st.C[0] = rng.standard_normal((3, 3))
    ns, cache = impl.step(st, _models()["nh"], CFG, advect=True)
    assert np.all(ns.v == 0) and np.all(ns.C == 0)
    assert np.array_equal(ns.F[0], F) and np.array_equal(ns.x[0], np.array(x, float))
    bars = [rng.standard_normal(s) for s in ((1, 3), (1, 3), (1, 3, 3), (1, 3, 3), (1, 3, 3))]
    res = impl.step_vjp(cache, *bars, advect=True)
    xb, vb, Cb, Fb, Fgb = res[:5]
    assert np.array_equal(xb, bars[0]) and np.array_equal(Fb, bars[3])
    assert np.array_equal(Fgb, bars[4])
    assert np.all(vb == 0) and np.all(Cb == 0)


@pytest.mark.parametrize("impl", [RL, RV])
def test_grid_mass_threshold(impl):
    """pmass = 1e-11: the stencil's low-weight nodes carry w pmass <= 1e-12 and
    are dropped, so v' = s v0 (sum of the weights above 0.1)."""
    mat = RL.Material(density=1e-11 / CFG.sp ** 3)
    cfg = RL.SimConfig(N=8, dx=0.1, dt=1e-3, damping=0.05, sp=0.02, material=mat)
    assert abs(cfg.pmass - 1e-11) < 1e-24
    x = np.array([0.413, 0.362, 0.447])
    v0 = np.array([0.3, -0.2, 0.5])
    X = x / cfg.dx
    S = 0.0
    for i in range(int(np.floor(X[0])) - 1, int(np.floor(X[0])) + 3):
        for j in range(int(np.floor(X[1])) - 1, int(np.floor(X[1])) + 3):
            for k in range(int(np.floor(X[2])) - 1, int(np.floor(X[2])) + 3):
                w = _N(i - X[0]) * _N(j - X[1]) * _N(k - X[2])
                assert abs(w * cfg.pmass - RL.MASS_EPS) > 1e-3 * RL.MASS_EPS   # no tie
                if w * cfg.pmass > RL.MASS_EPS:
                    S += w
    assert 0.2 < S < 0.9                                  # some nodes dropped, some kept
    ns, _ = impl.step(_single(x, v0), _models()["nh"], cfg, advect=True)
    assert np.allclose(ns.v[0], (1 - cfg.damping) * v0 * S, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("name", ["nh", "hencky", "hb"])
def test_edge_vjp_matches_central_differences(name):
    """A scene with particles straddling every face, one off the grid and the
    rest inside: <VJP(bar), d> vs the central difference of <bar, step(s + e d)>.
    No coordinate sits at a half-integer cell position, where the stencil base
    switches and dw/dx has a kink (one-sided limits differ there)."""
    rng = np.random.default_rng(11)
    pts = np.concatenate([np.array(STRADDLE), np.array([[-0.4, 0.3, 0.3]]),
                          rng.uniform(0.2, 0.6, (5, 3))])
    n = len(pts)
    st = RL.ParticleState(pts, 0.3 * rng.standard_normal((n, 3)),
                          0.5 * rng.standard_normal((n, 3, 3)),
                          I3 + 0.05 * rng.standard_normal((n, 3, 3)),
                          I3 * 1.1 + 0.03 * rng.standard_normal((n, 3, 3)))
    m = _models()[name]
    ns, cache = RV.step(st, m, CFG, advect=True)
    bars = [rng.standard_normal(a.shape) for a in (ns.x, ns.v, ns.C, ns.F, ns.Fg)]
    an = RV.step_vjp(cache, *bars, advect=True)
    d = [rng.standard_normal(a.shape) for a in (st.x, st.v, st.C, st.F, st.Fg)]
    eps = 1e-6

    def L(s):
        p = RL.ParticleState(*(a + s * eps * da for a, da in zip((st.x, st.v, st.C, st.F, st.Fg), d)))
        o, _ = RV.step(p, m, CFG, advect=True)
        return sum(float(np.sum(b * a)) for b, a in zip(bars, (o.x, o.v, o.C, o.F, o.Fg)))
    fd = (L(1) - L(-1)) / (2 * eps)
    dot = sum(float(np.sum(a * b)) for a, b in zip(an[:5], d))
    assert abs(dot - fd) <= 1e-6 * abs(fd), (dot, fd)


def test_planted_clamp_fails_the_edge_pin(monkeypatch):
    """An oracle that clamped off-grid nodes onto the border instead of skipping
    them must fail the straddling closed form."""
    real = RV._Stencil.__init__

    def clamped(self, x, cfg, need_grad):
        real(self, x, cfg, need_grad)
        Nx, Ny, Nz = cfg.shape
        nodes = self.base[:, None, :] + RV.OFFS[None]
        nodes = np.clip(nodes, 0, np.array([Nx, Ny, Nz]) - 1)
        self.flat = (nodes[..., 0] * Ny + nodes[..., 1]) * Nz + nodes[..., 2]
        self.inb = np.ones_like(self.inb)
    monkeypatch.setattr(RV._Stencil, "__init__", clamped)
    v0 = np.array([0.3, -0.2, 0.5])
    ns, _ = RV.step(_single(STRADDLE[0], v0), _models()["nh"], CFG, advect=True)
    monkeypatch.undo()
    vn, _ = _closed_form(np.array(STRADDLE[0]), v0, CFG)
    assert not np.allclose(ns.v[0], vn, rtol=1e-6)
