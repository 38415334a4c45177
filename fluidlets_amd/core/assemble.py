"""Dense kernel matrices on the GPU.

Drop-in for src/fluidlets/core/assemble.py: ``assemble`` (27-46),
``assemble_nearest`` (68-101) and ``nearest_neighbors`` (61-65).  Entry layout
follows the [point, component] convention: row p*i + a is component a at
target i, column q*j + b is component b of source j, so
``K @ forces.ravel()`` gives ``velocities.ravel()``.
"""
from __future__ import annotations

from contextlib import nullcontext

import numpy as np

from .. import _device as D
from .._lib import check, kernel_struct, lib
from ._common import blob_widths, field_dtype, points


def _eps_device(eps, n, dev):
    e = eps if not (n > 1 and np.all(eps == eps[0])) else eps[:1]
    return D.eps_arg(e, n, dev)


def assemble_device(kernel, tgt, src, eps_t, eps_stride, mu=1.0, fp64=True, fp32=False, stream=None):
    """Device assembly from (3, M) / (3, N) tensors.  Returns (K64, K32);
    either may be None.  For complex mu returns (Kc, None) with Kc complex128."""
    t = D.torch()
    M, N = int(tgt.shape[1]), int(src.shape[1])
    mu = complex(mu)
    k = kernel_struct(kernel)
    L = lib()
    if mu.imag != 0.0:
        Kc = t.empty((3 * M, 3 * N), dtype=t.complex128, device=tgt.device)
        st = L.fl_assemble(k, M, N, D.ptr(tgt), D.ptr(src), D.ptr(eps_t), eps_stride, mu.real, mu.imag,
                           None, 3 * N, None, 3 * N, D.ptr(Kc), D.stream_handle(stream))
        check(st, "fl_assemble")
        return Kc, None
    K64 = t.empty((3 * M, 3 * N), dtype=t.float64, device=tgt.device) if fp64 else None
    K32 = t.empty((3 * M, 3 * N), dtype=t.float32, device=tgt.device) if fp32 else None
    st = L.fl_assemble(k, M, N, D.ptr(tgt), D.ptr(src), D.ptr(eps_t), eps_stride, mu.real, 0.0,
                       D.ptr(K64), 3 * N, D.ptr(K32), 3 * N, None, D.stream_handle(stream))
    check(st, "fl_assemble")
    return K64, K32


PAD_MULTIPLE = 128   # dense solves run at orders rounded up to this


def padded_order(n: int, multiple: int = PAD_MULTIPLE) -> int:
    return -(-int(n) // multiple) * multiple


def assemble_padded_device(kernel, x_soa, eps_t, eps_stride, rhs, fp32: bool = True, stream=None,
                           multiple: int = PAD_MULTIPLE):
    """The resistance matrix of ``kernel`` on the (3, N) device points (mu =
    1) in the leading block of identity-padded buffers of order n_pad = 3N
    rounded up to ``multiple``, and the right-hand sides rhs (3N, q, float64,
    device) over zeros: [[K, 0], [0, I]] [x; 0] = [rhs; 0].  Partial pivoting
    never takes a padding row in the first 3N columns, so the factors are
    K's next to I and the solution is K's over zeros; the factorization of a
    ragged order is up to a quarter slower than at the next multiple of 128
    (misaligned rows and partial GEMM tiles).  Returns (K64, K32 or None,
    b (n_pad, q), n)."""
    t = D.torch()
    N = int(x_soa.shape[1])
    n = 3 * N
    q = int(rhs.shape[1])
    if N == 0 or int(rhs.shape[0]) != n or not 1 <= q <= 8:
        raise ValueError(f"{N} points do not match right-hand sides of shape {tuple(rhs.shape)}")
    n_pad = padded_order(n, multiple)
    dev = x_soa.device
    L = lib()
    with t.cuda.stream(stream) if stream is not None else nullcontext():
        K64 = t.empty((n_pad, n_pad), dtype=t.float64, device=dev)
        K32 = t.empty((n_pad, n_pad), dtype=t.float32, device=dev) if fp32 else None
        check(L.fl_assemble(kernel_struct(kernel), N, N, D.ptr(x_soa), D.ptr(x_soa), D.ptr(eps_t), eps_stride, 1.0,
                            0.0, D.ptr(K64), n_pad, D.ptr(K32), n_pad, None, D.stream_handle(stream)), "fl_assemble")
        n_dev = t.full((1,), float(n), dtype=t.float64, device=dev)
        src = rhs.contiguous()
        b = t.empty((n_pad, q), dtype=t.float64, device=dev)
        check(L.fl_pad_solve(D.ptr(K64), D.ptr(K32), n_pad, D.ptr(n_dev), D.ptr(src), D.ptr(b), q,
                             D.stream_handle(stream)), "fl_pad_solve")
    return K64, K32, b, n


def assemble_padded(kernel, points, eps, rhs, fp32: bool = True, stream=None, multiple: int = PAD_MULTIPLE):
    """``assemble_padded_device`` from host points (N, 3), blob widths and
    right-hand sides (3N, q)."""
    t = D.torch()
    dev = D.device()
    points = np.ascontiguousarray(points, dtype=np.float64)
    with t.cuda.stream(stream) if stream is not None else nullcontext():
        x_soa = D.soa(points, dev)
        eps_t, stride = D.eps_arg(eps, len(points), dev)
        rhs_t = t.from_numpy(np.ascontiguousarray(rhs, dtype=np.float64)).to(dev)
    return assemble_padded_device(kernel, x_soa, eps_t, stride, rhs_t, fp32=fp32, stream=stream, multiple=multiple)


def assemble(kernel, targets, sources, eps, mu=1.0) -> np.ndarray:
    """Dense (p M) x (q N) matrix of ``kernel`` from N sources to M targets
    (reference signature and result; a complex mu gives a complex matrix)."""
    kernel_struct(kernel)   # only the kernels the device implements (NotImplementedError otherwise)
    targets = points(targets, kernel.dim)
    sources = points(sources, kernel.dim)
    M, N = targets.shape[0], sources.shape[0]
    eps = blob_widths(eps, N)
    p, q = kernel.out_shape
    dtype = field_dtype(mu)
    if M == 0 or N == 0:
        return np.zeros((p * M, q * N), dtype=dtype)
    dev = D.device()
    eps_t, stride = _eps_device(eps, N, dev)
    K, _ = assemble_device(kernel, D.soa(targets, dev), D.soa(sources, dev), eps_t, stride, mu)
    out = K.cpu().numpy()
    return out if out.dtype == dtype else out.astype(dtype)


def nearest_groups(nearest: np.ndarray, n_force: int):
    """Quadrature indices grouped by force point, increasing within each group
    (the order in which core/assemble.py:53-58 accumulates them)."""
    nearest = np.ascontiguousarray(nearest, dtype=np.int64)
    if nearest.size and (nearest.min() < 0 or nearest.max() >= n_force):
        raise ValueError("nearest indices must lie in [0, n_force)")
    order = np.argsort(nearest, kind="stable").astype(np.int64)
    start = np.searchsorted(nearest[order], np.arange(n_force + 1)).astype(np.int64)
    return order, start


def assemble_nearest_device(kernel, tgt, quad, eps_t, eps_stride, order_t, start_t, n_force, mu=1.0,
                            K64=None, fp32=False, stream=None, fp64=True):
    t = D.torch()
    M, Q = int(tgt.shape[1]), int(quad.shape[1])
    if not (fp64 or fp32):
        raise ValueError("assemble_nearest_device needs fp64 or fp32 output")
    if K64 is None and fp64:
        K64 = t.empty((3 * M, 3 * n_force), dtype=t.float64, device=tgt.device)
    ld = int(K64.stride(0)) if K64 is not None else 3 * n_force
    K32 = t.empty((3 * M, 3 * n_force), dtype=t.float32, device=tgt.device) if fp32 else None
    st = lib().fl_assemble_nearest(kernel_struct(kernel), M, Q, int(n_force), D.ptr(tgt), D.ptr(quad), D.ptr(eps_t),
                                   eps_stride, D.ptr(order_t), D.ptr(start_t), float(mu), D.ptr(K64), ld,
                                   D.ptr(K32), 3 * n_force, D.stream_handle(stream))
    check(st, "fl_assemble_nearest")
    return K64, K32


def assemble_nearest(kernel, targets, quad_points, nearest, n_force, eps, mu=1.0) -> np.ndarray:
    """Kernel matrix with forces on a coarse set and quadrature on a fine set
    (Gallagher & Smith nearest-neighbour discretization):
    K[i, j] = sum over k with nearest[k] = j of G(x_i, X_k) / mu."""
    kernel_struct(kernel)   # only the kernels the device implements (NotImplementedError otherwise)
    targets = points(targets, kernel.dim)
    quad = points(quad_points, kernel.dim)
    Q = quad.shape[0]
    eps = blob_widths(eps, Q)
    nearest = np.ascontiguousarray(nearest, dtype=np.int64)
    p, q = kernel.out_shape
    n_force = int(n_force)
    dtype = field_dtype(mu)
    M = targets.shape[0]
    if M == 0 or n_force == 0:
        return np.zeros((p * M, q * n_force), dtype=dtype)
    order, start = nearest_groups(nearest, n_force)
    dev = D.device()
    t = D.torch()
    eps_t, stride = _eps_device(eps, Q, dev)
    mu_c = complex(mu)
    K, _ = assemble_nearest_device(kernel, D.soa(targets, dev), D.soa(quad, dev), eps_t, stride,
                                   t.from_numpy(order).to(dev), t.from_numpy(start).to(dev), n_force,
                                   mu=mu_c.real if mu_c.imag == 0 else 1.0)
    out = K.cpu().numpy()
    if mu_c.imag != 0:
        return out.astype(dtype) / mu
    return out.astype(dtype) if out.dtype != dtype else out


def nearest_neighbors(quad_points, force_points) -> np.ndarray:
    """Index of the nearest force point for every quadrature point, shape (Q,)."""
    from .. import _nearest
    return _nearest.nearest_neighbors(quad_points, force_points)
