"""Device GMRES for resistance problems (fl_gmres)."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _device as D
from .._lib import FL_ENOCONV, FL_FP32C, FL_FP64, check, kernel_struct, last_error, lib


def gmres_device(kernel, x_soa, eps_t, stride, U_soa, f_soa=None, mu=1.0, rtol=1e-10, restart=60, maxiter=200,
                 precision="fp64", stream=None, block_points: int = 0):
    """Solve on (3, N) device tensors; returns (f_soa, iterations, residual).
    Raises RuntimeError when GMRES does not converge (as the reference).
    block_points > 0 preconditions with the bodies' diagonal blocks (bodies
    of block_points consecutive points; FP64 only, fl_gmres_bj)."""
    t = D.torch()
    N = int(x_soa.shape[1])
    inner = FL_FP32C if precision in ("mixed", "fp32c") else FL_FP64
    if block_points and inner != FL_FP64:
        raise ValueError("block-Jacobi GMRES runs with FP64 matvecs")
    if f_soa is None:
        f_soa = t.zeros_like(U_soa)
    L = lib()
    its = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    if block_points:
        nbytes = int(L.fl_gmres_bj_workspace_size(N, int(restart), int(block_points)))
        ws = D.workspace(nbytes, stream)
        st = L.fl_gmres_bj(kernel_struct(kernel), N, D.ptr(x_soa), D.ptr(eps_t), int(stride), float(np.real(mu)),
                           D.ptr(U_soa), D.ptr(f_soa), float(rtol), int(restart), int(maxiter), int(block_points),
                           ctypes.byref(its), ctypes.byref(res), D.ptr(ws), nbytes, D.stream_handle(stream))
    else:
        nbytes = int(L.fl_gmres_workspace_size(N, int(restart), inner))
        ws = D.workspace(nbytes, stream)
        st = L.fl_gmres(kernel_struct(kernel), N, D.ptr(x_soa), D.ptr(eps_t), int(stride), float(np.real(mu)),
                        D.ptr(U_soa), D.ptr(f_soa), float(rtol), int(restart), int(maxiter), inner, ctypes.byref(its),
                        ctypes.byref(res), D.ptr(ws), nbytes, D.stream_handle(stream))
    if st == FL_ENOCONV:
        raise RuntimeError(f"GMRES did not converge ({last_error()})")
    check(st, "fl_gmres")
    return f_soa, int(its.value), float(res.value)


def gmres_solve(kernel, x_soa, eps_t, stride, rhs_flat, rtol=1e-10, restart=60, maxiter=200, precision="fp64",
                block_points: int = 0):
    """Host wrapper: rhs_flat is the (3N,) [point, component] right-hand side;
    returns the (3N,) solution in the same order and the iteration count."""
    N = int(x_soa.shape[1])
    dev = x_soa.device
    t = D.torch()
    U_soa = t.from_numpy(np.ascontiguousarray(np.asarray(rhs_flat, float).reshape(N, 3).T)).to(dev)
    f_soa, its, _ = gmres_device(kernel, x_soa, eps_t, stride, U_soa, rtol=rtol, restart=restart, maxiter=maxiter,
                                 precision=precision, block_points=block_points)
    return np.ascontiguousarray(f_soa.cpu().numpy().T).reshape(-1), its
