"""Kernel-agnostic assembly, matrix-free evaluation and solvers (device)."""
from .assemble import assemble, assemble_nearest, nearest_neighbors
from .matvec import evaluate
from .solve import DENSE_MAX, SolveInfo, grand_resistance, resistance

__all__ = ["assemble", "assemble_nearest", "nearest_neighbors", "evaluate", "resistance", "SolveInfo",
           "grand_resistance", "DENSE_MAX"]
