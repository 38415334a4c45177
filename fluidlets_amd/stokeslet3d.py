"""``stokeslet3d_velocity`` (src/fluidlets/stokeslet3d.py:15-32): the same as
``evaluate(Stokeslet3D(), targets, sources, forces, eps, mu)``."""
from __future__ import annotations

import numpy as np

from .core.matvec import evaluate
from .kernels import Stokeslet3D


def stokeslet3d_velocity(targets, sources, forces, eps, mu=1.0) -> np.ndarray:
    """Velocity at targets induced by 3D regularized Stokeslets, (M, 3)."""
    return evaluate(Stokeslet3D(), targets, sources, forces, eps, mu)
