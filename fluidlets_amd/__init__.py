"""fluidlets_amd: MI355X (gfx950) backend for the Fluidlets regularized-Stokeslet
hot path, with the reference's public API (src/fluidlets/__init__.py:2-17).

    import fluidlets_amd as fl
    u = fl.evaluate(fl.WallStokeslet3D(), targets, sources, forces, eps, mu)

Compute runs in hand-written HIP kernels (libfluidlets_hip.so, built in-tree
by ``python -m fluidlets_amd.build``) called through ctypes; PyTorch-ROCm
tensors hold device memory.  There is no CPU fallback.
"""
from .kernels import Brinkmanlet3D, Kernel, Stokeslet3D, VectorKernel, WallStokeslet3D
from .core import assemble, assemble_nearest, evaluate, grand_resistance, nearest_neighbors, resistance
from .stokeslet3d import stokeslet3d_velocity

__version__ = "0.1.0"
__all__ = [
    "stokeslet3d_velocity",
    "Stokeslet3D",
    "WallStokeslet3D",
    "Brinkmanlet3D",
    "assemble",
    "assemble_nearest",
    "nearest_neighbors",
    "evaluate",
    "resistance",
]
