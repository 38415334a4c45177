"""Matrix-free kernel sums u(x_i) = sum_j G(x_i, x_j) f_j / mu on the GPU.

Drop-in for src/fluidlets/core/matvec.py:24-49 (``evaluate``); the pair loop
(``_apply``, matvec.py:10-21) is the HIP kernel behind ``fl_eval``.
"""
from __future__ import annotations

import numpy as np

from .. import _device as D
from .._lib import FL_FP32C, FL_FP64, check, kernel_struct, lib
from ._common import blob_widths, field_dtype, points

PRECISIONS = {"fp64": FL_FP64, "fp32c": FL_FP32C}


def _precision(p) -> int:
    if isinstance(p, str):
        if p not in PRECISIONS:
            raise ValueError(f"unknown precision {p!r}")
        return PRECISIONS[p]
    return int(p)


def evaluate_device(kernel, tgt, src, eps_t, eps_stride, f_re, f_im=None, mu=1.0, u_re=None, u_im=None,
                    precision="fp64", stream=None):
    """Device-resident kernel sum.  tgt (3, M), src (3, N), f_re/f_im (3, N)
    float64 device tensors; returns (u_re, u_im) (3, M) tensors (u_im is None
    for a real result).  Enqueued on ``stream`` (default: torch's current)."""
    t = D.torch()
    M, N = int(tgt.shape[1]), int(src.shape[1])
    mu = complex(mu)
    complex_out = f_im is not None or mu.imag != 0.0
    if u_re is None:
        u_re = t.empty((3, M), dtype=t.float64, device=tgt.device)
    if complex_out and u_im is None:
        u_im = t.empty((3, M), dtype=t.float64, device=tgt.device)
    prec = _precision(precision)
    L = lib()
    nbytes = L.fl_eval_workspace_size(M, N, int(f_im is not None), prec)
    if M == N and int(eps_stride) == 0 and tgt.data_ptr() == src.data_ptr():
        # targets are the sources: room for the symmetric (reciprocal) evaluation
        nbytes = max(nbytes, L.fl_eval_sym_workspace_size(N, int(f_im is not None), prec))
    ws = D.workspace(nbytes, stream)
    k = kernel_struct(kernel)
    st = L.fl_eval(k, M, N, D.ptr(tgt), D.ptr(src), D.ptr(eps_t), int(eps_stride), D.ptr(f_re), D.ptr(f_im),
                   mu.real, mu.imag, D.ptr(u_re), D.ptr(u_im) if complex_out else None, prec,
                   D.ptr(ws), int(nbytes), D.stream_handle(stream))
    check(st, "fl_eval")
    return u_re, (u_im if complex_out else None)


def evaluate(kernel, targets, sources, strengths, eps, mu=1.0, precision="fp64") -> np.ndarray:
    """Field at targets from regularized sources, without storing the matrix.

    Same signature, dtypes, shapes and exceptions as the reference's
    ``fluidlets.evaluate`` (src/fluidlets/core/matvec.py:24-49); ``precision``
    ("fp64" default, or "fp32c") selects the device arithmetic.

    Returns (M, p) array, complex if mu or strengths are complex.
    """
    kernel_struct(kernel)   # only the kernels the device implements (NotImplementedError otherwise)
    targets = points(targets, kernel.dim)
    sources = points(sources, kernel.dim)
    N = sources.shape[0]
    eps = blob_widths(eps, N)
    p, q = kernel.out_shape
    dtype = field_dtype(mu, strengths)
    strengths = np.ascontiguousarray(np.asarray(strengths, dtype=dtype).reshape(N, q))
    M = targets.shape[0]
    if M == 0 or N == 0:
        return np.zeros((M, p), dtype=dtype) / mu
    dev = D.device()
    tgt = D.soa(targets, dev)
    same = targets is sources or (targets.shape == sources.shape and np.array_equal(targets, sources))
    src = tgt if same else D.soa(sources, dev)
    e = eps if not (N > 1 and np.all(eps == eps[0])) else eps[:1]
    eps_t, stride = D.eps_arg(e, N, dev)
    if np.iscomplexobj(strengths) and np.any(strengths.imag != 0):
        f_re, f_im = D.soa(strengths.real, dev), D.soa(strengths.imag, dev)
    elif np.iscomplexobj(strengths):
        f_re, f_im = D.soa(np.ascontiguousarray(strengths.real), dev), None
    else:
        f_re, f_im = D.soa(strengths, dev), None
    u_re, u_im = evaluate_device(kernel, tgt, src, eps_t, stride, f_re, f_im, mu, precision=precision)
    out = D.to_host_aos(u_re)
    if np.issubdtype(dtype, np.complexfloating):
        out = out + 1j * D.to_host_aos(u_im) if u_im is not None else out.astype(dtype)
    return out
