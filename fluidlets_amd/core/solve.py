"""Resistance problems on the device: velocities prescribed at points, forces
unknown.  Drop-in for src/fluidlets/core/solve.py:30-84.

``method="lu"`` assembles the dense matrix on the GPU and factors it with the
HIP LU (FP64, or ``precision="mixed"``: FP32 MFMA factors with FP64
iterative refinement); ``method="gmres"`` runs restarted GMRES on the device
with the matrix-free kernel sum as the operator.  As in the reference, a
final matrix-free residual ||K f - U|| / ||U|| is reported in SolveInfo.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .. import _device as D
from .. import linalg
from .._lib import kernel_struct
from ._common import blob_widths, field_dtype, points as as_points
from .assemble import assemble_padded_device
from .matvec import evaluate_device

#: above this many unknowns, method="auto" switches from dense LU to GMRES
DENSE_MAX = 6000


@dataclass
class SolveInfo:
    method: str
    iterations: int
    residual: float          # ||K f - U|| / ||U||
    precision: str = "fp64"
    refinement_steps: int = 0


def _eps_dev(eps, n, dev):
    e = eps if not (n > 1 and np.all(eps == eps[0])) else eps[:1]
    return D.eps_arg(e, n, dev)


def _rhs_columns(U: np.ndarray):
    """(3N,) real or complex host vector -> (3N, q) float64 columns."""
    if np.iscomplexobj(U):
        return np.ascontiguousarray(np.stack([U.real, U.imag], axis=1))
    return np.ascontiguousarray(U[:, None].astype(np.float64))


def _from_columns(x: np.ndarray, dtype, mu):
    f = x[:, 0] + 1j * x[:, 1] if x.shape[1] == 2 else x[:, 0]
    return (f * mu).astype(dtype) if np.issubdtype(dtype, np.complexfloating) else f * float(np.real(mu))


def symmetric_matrix(eps_stride: int) -> bool:
    """True when the kernel matrix is symmetric: every kernel here is
    reciprocal (G(y, x) = G(x, y)^T), so the matrix of one point set with one
    blob width is symmetric (per-source widths break it)."""
    return int(eps_stride) == 0


def lu_solve_device(kernel, x_soa, eps_t, stride, rhs, precision="fp64", tol=1e-10, max_iter=30, stream=None,
                    sym: bool | None = None):
    """Dense solve of K x = rhs (K assembled for mu = 1) on the device.
    rhs: (3N, q) float64 device tensor.  The system is factored at its order
    rounded up to a multiple of 128, identity-padded (assemble_padded_device:
    the same solution, a faster factorization).  A symmetric K (one blob
    width; ``sym`` None decides by symmetric_matrix) is factored as
    L D L^T (half the flops), with the verified LU as its fallback.  Returns
    (x (3N, q), refinement_steps, factor of the padded system)."""
    n = int(rhs.shape[0])
    if sym is None:
        sym = symmetric_matrix(stride)
    if precision == "fp64":
        def assemble64():
            return assemble_padded_device(kernel, x_soa, eps_t, stride, rhs, fp32=False, stream=stream)

        K64, _, b, _ = assemble64()
        F = linalg.getrf(K64, stream=stream, refill=lambda A: A.copy_(assemble64()[0]), sym=sym)
        if F.info > 0:
            raise np.linalg.LinAlgError("Matrix is singular.")
        linalg.getrs(F, b, stream=stream)
        return b[:n], 0, F
    if precision == "mixed":
        K64, K32, b, _ = assemble_padded_device(kernel, x_soa, eps_t, stride, rhs, fp32=True, stream=stream)
        F = linalg.getrf(K32, stream=stream, refill=lambda A: A.copy_(K64), products="bf16x3", sym=sym)
        if F.info > 0:
            raise np.linalg.LinAlgError("Matrix is singular.")
        x, steps, _ = linalg.refine(K64, F, b, tol=tol, max_iter=max_iter, stream=stream)
        F.extra["A64"] = K64
        return x[:n], steps, F
    raise ValueError(f"unknown precision {precision!r}")


def resistance(kernel, points, velocities, eps, mu=1.0, method="auto", rtol=1e-10, restart=60, maxiter=200,
               precision="fp64", block_points: int = 0):
    """Forces that produce the given velocities at the given points.

    Same signature, results and exceptions as the reference's
    ``fluidlets.resistance``; ``precision`` ("fp64" or "mixed") selects the
    dense factorization (LU) or the inner matvec precision (GMRES:
    "mixed" runs flexible GMRES with FP32C inner matvecs).  With GMRES,
    ``block_points`` > 0 (points per body, the bodies consecutive) adds
    block-Jacobi preconditioning by the bodies' diagonal blocks.

    Returns
    -------
    forces : (N, dim) array
    info : SolveInfo
    """
    kernel_struct(kernel)   # only the kernels the device implements (NotImplementedError otherwise)
    X = as_points(points, kernel.dim)
    n = X.shape[0]
    eps = blob_widths(eps, n)
    U = np.asarray(velocities).reshape(-1)
    dtype = field_dtype(mu, U)
    if method == "auto":
        method = "lu" if U.size <= DENSE_MAX else "gmres"
    if method not in ("lu", "gmres"):
        raise ValueError(f"unknown method {method!r}")
    t = D.torch()
    dev = D.device()
    x_soa = D.soa(X, dev)
    eps_t, stride = _eps_dev(eps, n, dev)
    cols = _rhs_columns(U)
    steps = 0
    if method == "lu":
        rhs = t.from_numpy(cols).to(dev)
        x, steps, _ = lu_solve_device(kernel, x_soa, eps_t, stride, rhs, precision=precision, tol=min(rtol, 1e-10))
        f = _from_columns(x.cpu().numpy(), dtype, mu)
        iterations = 1
    else:
        from .gmres import gmres_solve
        fcols = []
        iterations = 0
        for c in range(cols.shape[1]):
            fc, its = gmres_solve(kernel, x_soa, eps_t, stride, cols[:, c], rtol=rtol, restart=restart,
                                  maxiter=maxiter, precision=precision, block_points=block_points)
            fcols.append(fc)
            iterations += its
        f = _from_columns(np.stack(fcols, axis=1), dtype, mu)
    # final matrix-free residual, as core/solve.py:83
    fs = f.reshape(n, -1)
    if np.iscomplexobj(fs):
        fr, fi = D.soa(fs.real, dev), D.soa(fs.imag, dev)
    else:
        fr, fi = D.soa(fs, dev), None
    ur, ui = evaluate_device(kernel, x_soa, x_soa, eps_t, stride, fr, fi, mu)
    u = D.to_host_aos(ur)
    if ui is not None:
        u = u + 1j * D.to_host_aos(ui)
    resid = np.linalg.norm(u.ravel() - U) / np.linalg.norm(U)
    return fs, SolveInfo(method, iterations, float(resid), precision, steps)


def rigid_modes(X: np.ndarray, center) -> np.ndarray:
    """(3N, 6) velocities of the unit translations e_k and rotations
    e_k x (x - c), in [point, component] order."""
    r = np.asarray(X, float) - np.asarray(center, float)
    out = np.zeros((len(X), 3, 6))
    for k in range(3):
        out[:, k, k] = 1.0
        e = np.zeros(3)
        e[k] = 1.0
        out[:, :, 3 + k] = np.cross(e, r)
    return out.reshape(-1, 6)


def grand_resistance(kernel, points, center, eps, mu=1.0, precision="fp64", tol=1e-10):
    """6 x 6 grand resistance matrix of a rigid body: one factorization, six
    right-hand sides (unit translations and rotations about ``center``);
    column m is [sum f_m ; sum (x - c) x f_m] (the pattern of
    examples/translating_sphere/run.py:45-48 with all six modes at once).
    Returns (R, forces (6, N, 3), info dict)."""
    kernel_struct(kernel)
    X = as_points(points, 3)
    n = X.shape[0]
    eps = blob_widths(eps, n)
    t = D.torch()
    dev = D.device()
    x_soa = D.soa(X, dev)
    eps_t, stride = _eps_dev(eps, n, dev)
    rhs = t.from_numpy(rigid_modes(X, center)).to(dev)
    x, steps, F = lu_solve_device(kernel, x_soa, eps_t, stride, rhs, precision=precision, tol=tol)
    f = x.cpu().numpy().reshape(n, 3, 6).transpose(2, 0, 1) * float(np.real(mu))
    r = X - np.asarray(center, float)
    R = np.zeros((6, 6))
    for m in range(6):
        R[:3, m] = f[m].sum(axis=0)
        R[3:, m] = np.cross(r, f[m]).sum(axis=0)
    return R, f, {"refinement_steps": steps, "general_panels": F.general_panels}
