"""Seeded synthetic workloads (particle sets, configs, growth fields, schedules).

Workload choices the task statement leaves open are fixed here and recorded
in each builder's docstring.  Particles are placed 8 per cell (ppc = 8) as a
2x2x2 sub-cell lattice with a uniform jitter inside each sub-cell, so the
particle spacing (and SimConfig.sp) is dx/2.
"""
from dataclasses import dataclass, field

import numpy as np

from .config import Material, SimConfig


@dataclass
class Scene:
    name: str
    pts: np.ndarray            # (n, 3) float64 rest positions
    cfg: SimConfig
    model: object
    G: np.ndarray              # (n, 3, 3) growth tensor at the end of any ramp
    dG: np.ndarray             # (n, 3, 3) d G / d(growth parameter)
    n_steps: int
    sched: np.ndarray = None   # (n_steps,) a_t, Fg_t = I + a_t (G - I); None = static
    advect: bool = True
    checkpoint_every: int = None
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.pts)


def lattice(cells, dx, ppc=8, rng=None, jitter=True):
    """Particles in the given cells (int array (m, 3)): 2x2x2 sub-cells per
    cell (ppc must be 8), uniform jitter inside each sub-cell."""
    if ppc != 8:
        raise ValueError("only 8 particles per cell (2x2x2) are generated")
    cells = np.asarray(cells, dtype=np.float64)
    sub = np.array([(i, j, k) for i in range(2) for j in range(2) for k in range(2)], float)
    off = rng.uniform(0.0, 1.0, (len(cells), 8, 3)) if jitter else np.full((len(cells), 8, 3), 0.5)
    pts = (cells[:, None, :] + (sub[None] + off) * 0.5) * dx
    return pts.reshape(-1, 3)


def _box_cells(lo, hi):
    r = [np.arange(lo[a], hi[a]) for a in range(3)]
    g = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3)
    return g


def ramp(n_steps):
    """a_t = (t + 1) / n_steps: growth reaches G at the last step."""
    return (np.arange(n_steps, dtype=np.float64) + 1.0) / n_steps


def free_swell_cube(cells=8, origin=12, N=32, ppc=8, seed=0, rate=0.1, n_steps=100, dt=1e-4,
                    model="nh"):
    """Free swell of a cube (the CPU-runnable plain-method workload at its
    defaults): 32^3 grid, dx = 1/32, cells [12, 20)^3, 8 ppc -> 4096 particles,
    neo-Hookean E = 1e3, nu = 0.3, rho = 1, damping 0.05 (reference default),
    isotropic growth G = (1 + rate) I ramped linearly over n_steps, dt = 1e-4."""
    from . import constitutive as K
    rng = np.random.default_rng(seed)
    dx = 1.0 / N
    pts = lattice(_box_cells([origin] * 3, [origin + cells] * 3), dx, ppc, rng)
    mat = Material(E=1.0e3, nu=0.3, density=1.0)
    cfg = SimConfig(N=N, dx=dx, dt=dt, damping=0.05, sp=dx / 2, material=mat)
    m = {"nh": K.NeoHookean, "hencky": K.Hencky}[model](mat)
    n = len(pts)
    eye = np.repeat(np.eye(3)[None], n, 0)
    return Scene("free_swell", pts, cfg, m, (1.0 + rate) * eye, eye.copy(), n_steps,
                 sched=ramp(n_steps), advect=True, extra={"rate": rate})


def bilayer_strip(ratio=1.08, n_steps=500, seed=0, dims=(96, 32, 32), lo=(16, 8, 12),
                  hi=(80, 24, 20), dt=5e-5, model="hencky"):
    """Bilayer strip: 96 x 32 x 32 grid, dx = 1/96, strip cells x in [16, 80),
    y in [8, 24), z in [12, 20) -> 65,536 particles; top half (y >= 16 cells)
    grows 1.00, bottom half grows `ratio`; Hencky with the free-swell material;
    static growth, dt = 5e-5, 500 steps.  Growth parameter: the ratio."""
    from . import constitutive as K
    rng = np.random.default_rng(seed)
    dx = 1.0 / dims[0]
    cells = _box_cells(lo, hi)
    pts = lattice(cells, dx, 8, rng)
    ymid = 0.5 * (lo[1] + hi[1]) * dx
    bottom = (pts[:, 1] < ymid).astype(np.float64)
    mat = Material(E=1.0e3, nu=0.3, density=1.0)
    cfg = SimConfig(N=dims[0], dx=dx, dt=dt, damping=0.05, sp=dx / 2, material=mat, dims=dims)
    m = {"nh": K.NeoHookean, "hencky": K.Hencky}[model](mat)
    eye = np.eye(3)
    G = (1.0 + (ratio - 1.0) * bottom)[:, None, None] * eye
    dG = bottom[:, None, None] * eye
    tip = pts[:, 0] >= (hi[0] - 1) * dx
    h = (hi[1] - lo[1]) * dx
    return Scene("bilayer", pts, cfg, m, G, dG, n_steps, sched=None, advect=True,
                 extra={"ratio": ratio, "bottom": bottom, "tip": tip, "thickness": h,
                        "length": (hi[0] - lo[0]) * dx})


def batch_instances(count=256, seed=0, n_lo=1000, n_hi=32768, rate_lo=0.02, rate_hi=0.2,
                    n_steps=200, dx=1.0 / 64, dt=1e-4, margin=4, model="mixed"):
    """Batched inverse-growth workload: `count` instances, particle counts
    n_i = 8 ceil(exp(U(ln n_lo, ln n_hi)) / 8), growth rates U(rate_lo,
    rate_hi), one rng (seed 0).  Even i: free-swell cube; odd i: bilayer strip
    (aspect 4 : 1 : 1/2, bottom layer grows).  Cells are filled in lexicographic
    order until n_i / 8 cells; grid = block + `margin` nodes per side
    (4 recorded here; the stencil needs 2).  Free-swell material, dx = 1/64,
    dt = 1e-4, growth ramped over n_steps.  ``model``: "mixed" (the default:
    neo-Hookean free swells, Hencky bilayers, as the free-swell and bilayer
    workloads use them), "nh" or "hencky" for every instance."""
    from . import constitutive as K
    rng = np.random.default_rng(seed)
    ns = 8 * np.ceil(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count)) / 8).astype(int)
    rates = rng.uniform(rate_lo, rate_hi, count)
    mat = Material(E=1.0e3, nu=0.3, density=1.0)
    scenes = []
    for i in range(count):
        ncell = ns[i] // 8
        if i % 2 == 0:
            a = int(np.ceil(ncell ** (1 / 3)))
            bd = (a, a, a)
        else:
            a = int(np.ceil((ncell / 2.0) ** (1 / 3)))       # (4a, a, a/2)
            bd = (4 * a, a, max(1, (a + 1) // 2))
            while bd[0] * bd[1] * bd[2] < ncell:
                bd = (bd[0] + 1, bd[1], bd[2])
        cells = _box_cells([margin] * 3, [margin + b for b in bd])[:ncell]
        pts = lattice(cells, dx, 8, rng)
        dims = tuple(int(b + 2 * margin) for b in bd)
        cfg = SimConfig(N=max(dims), dx=dx, dt=dt, damping=0.05, sp=dx / 2, material=mat,
                        dims=dims)
        n = len(pts)
        eye = np.repeat(np.eye(3)[None], n, 0)
        if i % 2 == 0:
            dG = eye.copy()
        else:
            ymid = (margin + 0.5 * bd[1]) * dx
            dG = (pts[:, 1] < ymid).astype(float)[:, None, None] * np.eye(3)
        G = eye + rates[i] * dG
        kind = model if model != "mixed" else ("nh" if i % 2 == 0 else "hencky")
        scenes.append(Scene("free_swell" if i % 2 == 0 else "bilayer", pts, cfg,
                            K.NeoHookean(mat) if kind == "nh" else K.Hencky(mat), G, dG, n_steps,
                            sched=ramp(n_steps), advect=True, checkpoint_every=20,
                            extra={"rate": rates[i]}))
    return scenes


def bioink_blob(cells=64, N=128, seed=0, n_steps=300, dt=1e-4, tau_y=50.0, K=10.0, n=0.6):
    """Rate-dependent bioink: 128^3 grid, dx = 1/128, centred 64^3-cell blob at
    8 ppc -> 2,097,152 particles; Herschel-Bulkley tau_y = 50, K = 10,
    n = 0.6, m = 100, eps = 1e-4 (reference defaults), with kappa = the bulk
    modulus lambda + 2 mu / 3 of the reference's default material (E = 1e4,
    nu = 0.3, rho = 1000, config.py:7-11); radially graded static growth
    g(r) = 1 + 0.1 (1 - r), r = distance from the centre / blob half-width
    clipped to [0, 1]; dt = 1e-4, 300 steps, checkpoint every 25.  (At
    rho = 1 the regularised viscosity tau_y m ~ 5e3 makes explicit steps at
    this dx diverge within a few steps.)"""
    from . import constitutive as Km
    rng = np.random.default_rng(seed)
    dx = 1.0 / N
    lo = (N - cells) // 2
    pts = lattice(_box_cells([lo] * 3, [lo + cells] * 3), dx, 8, rng)
    mat = Material(E=1.0e4, nu=0.3, density=1000.0)
    kappa = mat.lam + 2.0 * mat.mu / 3.0
    cfg = SimConfig(N=N, dx=dx, dt=dt, damping=0.05, sp=dx / 2, material=mat)
    model = Km.HerschelBulkley(kappa=kappa, K=K, n=n, tau_y=tau_y, m=100.0, eps=1e-4)
    c = 0.5 * N * dx
    r = np.clip(np.linalg.norm(pts - c, axis=1) / (0.5 * cells * dx), 0.0, 1.0)
    g = 1.0 + 0.1 * (1.0 - r)
    eye = np.eye(3)
    G = g[:, None, None] * eye
    return Scene("bioink", pts, cfg, model, G, np.zeros_like(G), n_steps, sched=None,
                 advect=True, checkpoint_every=25)


def shell(N=100, r_in=0.25, r_out=0.35, ctrl=16, seed=0, n_steps=300, dt=5e-5, std=0.01):
    """Heterogeneous growth field on a spherical shell: 100^3 cells, dx = 0.01,
    shell centred at 0.5 with radii 0.25 / 0.35, 8 ppc in cells whose centre is
    in the shell (~0.9 M particles); Fg = I + diag(a, b, c) trilinearly
    interpolated at rest positions from a 16^3 x 3 control grid over [0, 1]^3,
    control values N(0, std^2) with std = 0.01; Hencky, 300 steps."""
    from . import constitutive as K
    rng = np.random.default_rng(seed)
    dx = 1.0 / N
    cells = _box_cells([0, 0, 0], [N, N, N])
    cc = (cells + 0.5) * dx - 0.5
    rr = np.linalg.norm(cc, axis=1)
    cells = cells[(rr >= r_in) & (rr <= r_out)]
    pts = lattice(cells, dx, 8, rng)
    mat = Material(E=1.0e3, nu=0.3, density=1.0)
    cfg = SimConfig(N=N, dx=dx, dt=dt, damping=0.05, sp=dx / 2, material=mat)
    theta = rng.normal(0.0, std, (ctrl, ctrl, ctrl, 3))
    W, idx = trilinear_weights(pts, ctrl)
    diag = np.einsum("pk,pkc->pc", W, theta.reshape(-1, 3)[idx])
    G = np.repeat(np.eye(3)[None], len(pts), 0)
    G[:, [0, 1, 2], [0, 1, 2]] += diag
    return Scene("shell", pts, cfg, K.Hencky(mat), G, np.zeros_like(G), n_steps, sched=None,
                 advect=True, checkpoint_every=25,
                 extra={"theta": theta, "W": W, "idx": idx, "ctrl": ctrl})


def trilinear_weights(pts, ctrl):
    """Weights (n, 8) and flat control indices (n, 8) of trilinear
    interpolation on a ctrl^3 node grid spanning [0, 1]^3."""
    u = np.clip(pts, 0.0, 1.0) * (ctrl - 1)
    i0 = np.minimum(np.floor(u).astype(np.int64), ctrl - 2)
    f = u - i0
    W = np.empty((len(pts), 8))
    idx = np.empty((len(pts), 8), dtype=np.int64)
    k = 0
    for a in range(2):
        for b in range(2):
            for c in range(2):
                W[:, k] = ((f[:, 0] if a else 1 - f[:, 0]) * (f[:, 1] if b else 1 - f[:, 1])
                           * (f[:, 2] if c else 1 - f[:, 2]))
                idx[:, k] = ((i0[:, 0] + a) * ctrl + (i0[:, 1] + b)) * ctrl + (i0[:, 2] + c)
                k += 1
    return W, idx


def timoshenko_bilayer(eps, axial=False, res=1, n_steps=4000):
    """The reference's bilayer-bending scene (tests/test_growth.cpp:242-263):
    domain 1.2 (48 res cells, dx = 0.025 / res), beam L = 0.5, h = 0.05, depth
    0.025 centred at 0.6, particles on a lattice of spacing dx / 2 (the
    reference accumulates the coordinates in float32, reproduced here), E = 1e4,
    nu = 0.3, rho = 1000, damping 0.05, dt = 1e-3; the top half grows 1 + eps
    (isotropically, or along x only with ``axial``), the bottom half not.
    Hencky.  extra["kappa_timoshenko"] = 1.5 eps / h."""
    from . import constitutive as K
    sp = np.float32(0.0125 / res)
    c = np.float32(0.6)

    def axis(half):
        v, out = np.float32(c - np.float32(half)), []
        while v <= np.float32(c + np.float32(half) + np.float32(1e-6)):
            out.append(float(v))
            v = np.float32(v + sp)
        return np.array(out)
    xs, ys, zs = axis(0.25), axis(0.025), axis(0.0125)
    pts = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), -1).reshape(-1, 3)
    mat = Material(E=1.0e4, nu=0.3, density=1000.0)
    cfg = SimConfig(N=48 * res, dx=0.025 / res, dt=1.0e-3, damping=0.05, sp=float(sp),
                    material=mat)
    n = len(pts)
    top = pts[:, 1] >= float(c)
    g = np.where(top, 1.0 + eps, 1.0)
    G = np.zeros((n, 3, 3))
    G[:, 0, 0] = g
    G[:, 1, 1] = 1.0 if axial else g
    G[:, 2, 2] = 1.0 if axial else g
    dG = np.zeros((n, 3, 3))
    dG[:, 0, 0] = top
    if not axial:
        dG[:, 1, 1] = top
        dG[:, 2, 2] = top
    return Scene("timoshenko_bilayer", pts, cfg, K.Hencky(mat), G, dG, n_steps, sched=None,
                 advect=True, extra={"eps": eps, "kappa_timoshenko": 1.5 * eps / 0.05})

