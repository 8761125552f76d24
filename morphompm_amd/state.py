"""Particle state (python/morphompm/state.py:10-31).

Fields are (n, 3) / (n, 3, 3) arrays, numpy or torch.  The GPU path packs them
into the structure-of-arrays slot layout of include/morphompm_hip.h and
unpacks results into the caller's array kind, dtype and (for torch) device.
"""
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class ParticleState:
    x: object
    v: object
    C: object
    F: object
    Fg: object

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @staticmethod
    def rest(x0) -> "ParticleState":
        """Particles at rest: v = C = 0, F = Fg = I."""
        n = x0.shape[0]
        if isinstance(x0, torch.Tensor):
            eye = torch.eye(3, dtype=x0.dtype, device=x0.device).expand(n, 3, 3)
            return ParticleState(x0.clone(), torch.zeros_like(x0),
                                 torch.zeros(n, 3, 3, dtype=x0.dtype, device=x0.device),
                                 eye.clone(), eye.clone())
        x0 = np.asarray(x0, dtype=np.float64)
        eye = np.repeat(np.eye(3)[None], n, axis=0)
        return ParticleState(x0.copy(), np.zeros((n, 3)), np.zeros((n, 3, 3)), eye.copy(),
                             eye.copy())
