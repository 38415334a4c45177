"""Kernel-agnostic assembly, matrix-free evaluation and solvers (device)."""
from .assemble import assemble, assemble_nearest, nearest_neighbors
from .matvec import evaluate

__all__ = ["assemble", "assemble_nearest", "nearest_neighbors", "evaluate"]
