"""Material and discretisation (python/morphompm/config.py:7-40).

Same fields, defaults and derived quantities as the reference.  ``dims`` is an
extension: a non-cubic node grid (Nx, Ny, Nz); None keeps the cubic N^3 grid.
"""
from dataclasses import dataclass, field


@dataclass
class Material:
    E: float = 1.0e4
    nu: float = 0.3
    density: float = 1000.0

    @property
    def mu(self) -> float:
        return self.E / (2.0 * (1.0 + self.nu))

    @property
    def lam(self) -> float:
        return self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))


@dataclass
class SimConfig:
    N: int = 8
    dx: float = 0.1
    dt: float = 1.0e-3
    damping: float = 0.05
    sp: float = 0.02
    material: Material = field(default_factory=Material)
    dims: tuple = None

    @property
    def inv_dx(self) -> float:
        return 1.0 / self.dx

    @property
    def pmass(self) -> float:
        return self.material.density * self.sp ** 3

    @property
    def vol0(self) -> float:
        return self.sp ** 3

    @property
    def shape(self) -> tuple:
        return tuple(int(d) for d in self.dims) if self.dims is not None else (self.N,) * 3
