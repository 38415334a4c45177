"""Dense LU factorizations and solves on the device (fl_?getrf, fl_?getrs,
fl_refine).  Device memory is torch tensors; the arithmetic is the HIP library.
"""
from __future__ import annotations

import ctypes
from contextlib import nullcontext
from dataclasses import dataclass, field

import numpy as np

from . import _device as D
from ._lib import FL_FP32C, FL_FP64, FL_LU_BF16X3, FL_LU_FP32, check, lib


@dataclass
class LUFactor:
    """Factors of P A = L U kept on the device, with what the solves need.

    ``lu`` (n x n, float32 or float64) holds L (unit lower) and U; ``ipiv`` the
    LAPACK interchanges; ``aux`` the row permutation and the inverted diagonal
    blocks; ``info`` > 0 flags an exactly zero pivot (1-based column);
    ``general_panels`` counts panels that needed the all-rows pivot search."""
    lu: object
    ipiv: object
    aux: object
    n: int
    info: int = 0
    general_panels: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def precision(self) -> int:
        return FL_FP64 if str(self.lu.dtype) == "torch.float64" else FL_FP32C

    def nbytes(self) -> int:
        return int(self.lu.numel() * self.lu.element_size() + self.aux.numel())


def getrf(A, overwrite: bool = True, stream=None, refill=None, hint: dict | None = None,
          products: str = "fp32", sym: bool = False) -> LUFactor:
    """Factor an (n, n) float32/float64 device tensor in place (or a copy).

    ``products`` ("fp32" or "bf16x3", float32 only) selects the arithmetic of
    the large trailing updates (fl_sgetrf_ex): "bf16x3" splits every operand
    into two bf16 parts and runs three bf16 MFMA products, about twice as
    fast, with products good to about 2^-16 relative: for factors that FP64
    refinement (``refine``) uses.

    With ``refill`` (a callable that rewrites A's original values into the
    tensor it is given) the factorization first runs asynchronously without
    interchanges (fl_?getrf_async, one read-back for the whole matrix); if
    that is not partial pivoting's answer, A is refilled and factored on the
    verified path.  The result is the same either way.  ``hint`` (a dict
    shared by a sequence of similar factorizations) records a failed attempt,
    after which the sequence goes straight to the verified path.

    ``sym`` (with ``refill``): A is symmetric (a kernel matrix with one blob
    width) and the first attempt is the symmetric factorization A = L D L^T
    (fl_?sytrf_async: half the flops, the same factor format); a matrix whose
    multipliers exceed 1 is refilled and factored on the verified LU path."""
    t = D.torch()
    if A.dim() != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("getrf needs a square matrix")
    if A.dtype not in (t.float32, t.float64):
        raise ValueError("getrf needs float32 or float64")
    if not overwrite:
        A = A.clone()
    if not A.is_contiguous():
        A = A.contiguous()
    n = int(A.shape[0])
    prec = FL_FP64 if A.dtype == t.float64 else FL_FP32C
    if products not in ("fp32", "bf16x3"):
        raise ValueError(f"unknown products {products!r}")
    if sym and refill is None:
        raise ValueError("the symmetric factorization needs refill (its fallback is the verified LU)")
    split = FL_LU_BF16X3 if products == "bf16x3" and prec == FL_FP32C else FL_LU_FP32
    L = lib()
    ipiv = t.empty((max(n, 1),), dtype=t.int32, device=A.device)
    aux = t.empty((max(int(L.fl_getrf_aux_size(n, prec)), 1),), dtype=t.uint8, device=A.device)
    nbytes = int(L.fl_getrf_workspace_size(n, prec))
    ws = D.workspace(nbytes, stream)
    if refill is not None and (hint is None or hint.get("async", True)):
        status = t.empty((1,), dtype=t.int32, device=A.device)
        if prec == FL_FP64 and sym:
            check(L.fl_dsytrf_async(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), D.ptr(status), D.ptr(ws), nbytes,
                                    D.stream_handle(stream)), "sytrf_async")
        elif prec == FL_FP64:
            check(L.fl_dgetrf_async(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), D.ptr(status), D.ptr(ws), nbytes,
                                    D.stream_handle(stream)), "getrf_async")
        elif sym:
            check(L.fl_ssytrf_async(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), D.ptr(status), split, D.ptr(ws),
                                    nbytes, D.stream_handle(stream)), "sytrf_async")
        else:
            check(L.fl_sgetrf_async_ex(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), D.ptr(status), split, D.ptr(ws),
                                       nbytes, D.stream_handle(stream)), "getrf_async")
        with t.cuda.stream(stream) if stream is not None else nullcontext():
            ok = int(status.cpu().item()) == 0     # waits for the factorization on its stream
            if ok:
                return LUFactor(A, ipiv, aux, n, 0, 0, {"async": True, "products": products, "sym": bool(sym)})
            if hint is not None:
                hint["async"] = False
            refill(A)
    info = ctypes.c_int(0)
    fb = ctypes.c_int(0)
    if prec == FL_FP64:
        check(L.fl_dgetrf(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), ctypes.byref(info), ctypes.byref(fb), D.ptr(ws),
                          nbytes, D.stream_handle(stream)), "getrf")
    else:
        check(L.fl_sgetrf_ex(D.ptr(A), n, n, D.ptr(ipiv), D.ptr(aux), ctypes.byref(info), ctypes.byref(fb), split,
                             D.ptr(ws), nbytes, D.stream_handle(stream)), "getrf")
    return LUFactor(A, ipiv, aux, n, int(info.value), int(fb.value), {"products": products})


def getrs(F: LUFactor, B, stream=None):
    """Solve A X = B in place; B is an (n, nrhs) or (n,) device tensor of the
    factor's dtype.  Right-hand sides are processed 8 at a time."""
    L = lib()
    vec = B.dim() == 1
    B2 = B.view(-1, 1) if vec else B
    if not B2.is_contiguous():
        raise ValueError("B must be contiguous")
    n, nrhs = int(B2.shape[0]), int(B2.shape[1])
    if n != F.n:
        raise ValueError("B has the wrong number of rows")
    if B2.dtype != F.lu.dtype:
        raise ValueError("B must have the factor's dtype")
    prec = F.precision
    fn = L.fl_dgetrs if prec == FL_FP64 else L.fl_sgetrs
    for c0 in range(0, nrhs, 8):
        q = min(8, nrhs - c0)
        nbytes = int(L.fl_getrs_workspace_size(n, q, prec))
        ws = D.workspace(nbytes, stream)
        blk = B2[:, c0:c0 + q]
        check(fn(D.ptr(F.lu), n, n, D.ptr(F.aux), blk.data_ptr(), q, nrhs, D.ptr(ws), nbytes,
                 D.stream_handle(stream)), "getrs")
    return B


def getri(F: LUFactor, stream=None):
    """Explicit inverse (n x n float32 device tensor) from FP32 factors
    (fl_sgetri): for many refinements against one factorization."""
    t = D.torch()
    if F.precision != FL_FP32C:
        raise ValueError("getri needs FP32 factors")
    n = F.n
    X = t.empty((n, n), dtype=t.float32, device=F.lu.device)
    check(lib().fl_sgetri(D.ptr(F.lu), n, n, D.ptr(F.aux), D.ptr(X), n, D.stream_handle(stream)), "getri")
    return X


def refine_steps(A64, F32: LUFactor, b, x, resid, steps: int, ws, stream=None):
    """Exactly ``steps`` refinement corrections of A64 x = b from the FP32
    factors (fl_refine_steps), launched on ``stream`` with no host
    synchronization and no allocation, so that the call can be captured in a
    HIP graph.  b, x: (n, nrhs <= 8) float64 device tensors; ``resid``
    (float64, at least nrhs) receives the final relative residuals on the
    device; ``ws`` is a uint8 device tensor of fl_refine_workspace_size bytes.
    For the same factors x is bit for bit ``refine``'s result when ``refine``
    stops after ``steps`` corrections."""
    L = lib()
    n, nrhs = int(b.shape[0]), int(b.shape[1])
    if not (b.is_contiguous() and x.is_contiguous()) or tuple(x.shape) != (n, nrhs) or n != F32.n:
        raise ValueError("refine_steps needs contiguous (n, nrhs) b and x matching the factors")
    if F32.precision != FL_FP32C or int(A64.shape[0]) != n or resid.numel() < nrhs:
        raise ValueError("refine_steps needs FP32 factors of A64 and room for nrhs residuals")
    nbytes = int(L.fl_refine_workspace_size(n, nrhs))
    if ws.numel() < nbytes:
        raise ValueError(f"refine_steps needs {nbytes} workspace bytes")
    check(L.fl_refine_steps(D.ptr(A64), int(A64.stride(0)), D.ptr(F32.lu), int(F32.lu.stride(0)), D.ptr(F32.aux), n,
                            D.ptr(b), D.ptr(x), nrhs, nrhs, int(steps), D.ptr(resid), D.ptr(ws), nbytes,
                            D.stream_handle(stream)), "refine_steps")
    return x


def refine(A64, F32: LUFactor | None, b, tol: float = 1e-10, max_iter: int = 30, stream=None, frame=None,
           inverse=None):
    """x solving A64 x = b (FP64) from the FP32 factors by iterative
    refinement with FP64 residuals.  b: (n, nrhs) float64 device tensor.
    With `frame` (3x3, lab = frame @ body) the factors are those of the
    operator written in the body frame (A64 ~ R A' R^T per point), and every
    correction is rotated accordingly (fl_refine_frame).  With `inverse` (an
    explicit FP32 inverse from ``getri``; F32 may then be None) each
    correction is one product with it (fl_refine_inv).
    Returns (x, iterations, per-column final relative residuals)."""
    t = D.torch()
    L = lib()
    b2 = b.view(-1, 1) if b.dim() == 1 else b
    b2 = b2.contiguous()
    n, nrhs = int(b2.shape[0]), int(b2.shape[1])
    x = t.empty_like(b2)
    iters, resid = [], []
    for c0 in range(0, nrhs, 8):
        q = min(8, nrhs - c0)
        bb = b2[:, c0:c0 + q].contiguous()
        xx = t.empty_like(bb)
        nbytes = int(L.fl_refine_workspace_size(n, q))
        ws = D.workspace(nbytes, stream)
        it = ctypes.c_int(0)
        res = (ctypes.c_double * 8)()
        if inverse is not None:
            R = None if frame is None else np.ascontiguousarray(np.asarray(frame, dtype=np.float64).reshape(3, 3))
            st = L.fl_refine_inv(D.ptr(A64), n, D.ptr(inverse), int(inverse.stride(0)), n,
                                 None if R is None else R.ctypes.data, D.ptr(bb), D.ptr(xx), q, q, float(tol),
                                 int(max_iter), ctypes.byref(it), res, D.ptr(ws), nbytes, D.stream_handle(stream))
        elif frame is None:
            st = L.fl_refine(D.ptr(A64), n, D.ptr(F32.lu), n, D.ptr(F32.aux), n, D.ptr(bb), D.ptr(xx), q, q,
                             float(tol), int(max_iter), ctypes.byref(it), res, D.ptr(ws), nbytes,
                             D.stream_handle(stream))
        else:
            R = np.ascontiguousarray(np.asarray(frame, dtype=np.float64).reshape(3, 3))
            st = L.fl_refine_frame(D.ptr(A64), n, D.ptr(F32.lu), n, D.ptr(F32.aux), n,
                                   R.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), D.ptr(bb), D.ptr(xx), q, q,
                                   float(tol), int(max_iter), ctypes.byref(it), res, D.ptr(ws), nbytes,
                                   D.stream_handle(stream))
        check(st, "refine")
        x[:, c0:c0 + q] = xx
        iters.append(int(it.value))
        resid.extend(res[:q])
    return (x.view(-1) if b.dim() == 1 else x), max(iters), np.array(resid)
