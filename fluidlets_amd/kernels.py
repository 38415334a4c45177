"""Kernel descriptors: the same classes and constructors as the reference
(src/fluidlets/kernels/base.py:14-32, stokeslet3d.py:53-64, images.py:92-104,
brinkmanlet3d.py:107-141).

In the reference a kernel carries a Numba ``block`` function that the core
loops call per pair.  Here a kernel carries ``kind`` (and ``alpha``), which
selects the fused HIP pair kernel in libfluidlets_hip.so; the per-pair
``velocity`` method runs through the same device path.
"""
from __future__ import annotations

import numpy as np

STOKESLET3D, WALL_STOKESLET3D, BRINKMANLET3D = 0, 1, 2


class Kernel:
    """Base class.  Subclasses set ``dim``, ``out_shape`` and ``kind``."""

    dim: int
    out_shape: tuple[int, int]
    kind: int
    alpha: float = 0.0
    #: the reference's per-pair Numba function; the device kernels replace it
    block = None

    def velocity(self, x, x0, f, eps, mu=1.0) -> np.ndarray:
        """Field at one target from one source (one pair on the device)."""
        from .core.matvec import evaluate
        return evaluate(self, np.asarray(x, float)[None, :], np.asarray(x0, float)[None, :],
                        np.asarray(f)[None, :], eps, mu)[0]


class VectorKernel(Kernel):
    """Fluid kernel: a force (dim) at a source gives a velocity (dim) at a target."""

    @property
    def out_shape(self) -> tuple[int, int]:
        return (self.dim, self.dim)


class Stokeslet3D(VectorKernel):
    """Regularized Stokeslet in free space (Cortez, Fauci & Medovikov 2005)."""

    dim = 3
    kind = STOKESLET3D

    def __repr__(self) -> str:
        return "Stokeslet3D()"


class WallStokeslet3D(VectorKernel):
    """Regularized Stokeslet above a no-slip wall at z = 0 (sources need z > 0),
    with the exact image system of Ainley et al. (2008)."""

    dim = 3
    kind = WALL_STOKESLET3D

    def __repr__(self) -> str:
        return "WallStokeslet3D()"


class Brinkmanlet3D(VectorKernel):
    """Regularized Brinkmanlet in free space, with inverse screening length alpha >= 0.

    ``Brinkmanlet3D(0.0)`` is the regularized Stokeslet."""

    dim = 3
    kind = BRINKMANLET3D

    def __init__(self, alpha: float):
        if np.iscomplexobj(alpha) or alpha < 0:
            raise ValueError("alpha must be real and non-negative")
        self.alpha = float(alpha)

    def __repr__(self) -> str:
        return f"Brinkmanlet3D(alpha={self.alpha})"
