"""MI355X (gfx950) MLS-MPM growth step + hand-written adjoint.

Drop-in for morphompm's seams: ``constitutive`` (stress / stress_vjp),
``transfer`` (step / step_vjp), ``integrate`` (rollout / rollout_vjp /
dloss_dg), ``diff`` (infer_growth).  All compute runs in the HIP library
libmorphompm_hip.so through the C ABI in include/morphompm_hip.h.
"""
from . import config, state  # noqa: F401
from .config import Material, SimConfig  # noqa: F401
from .state import ParticleState  # noqa: F401

__version__ = "0.1.0"
