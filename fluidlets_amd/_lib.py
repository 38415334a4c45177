"""ctypes binding of libfluidlets_hip.so (the C ABI in include/fluidlets_hip.h).

The library is built in-tree (``python -m fluidlets_amd.build``).  There is
no fallback: if the library is missing, importing a compute function raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfluidlets_hip.so"

FL_OK, FL_EINVAL, FL_EHIP, FL_ENOCONV, FL_ESINGULAR, FL_EWORKSPACE = 0, -1, -2, -3, -4, -5
FL_FP64, FL_FP32C = 0, 1
FL_LU_FP32, FL_LU_BF16X3 = 0, 1
FL_EVAL_NOALLOC_SCRATCH_BYTES = 32 << 20


class fl_kernel(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("reserved", c_int32), ("alpha", c_double)]


_SIGNATURES = {
    "fl_version": (c_int, []),
    "fl_get_last_error": (c_int, [ctypes.c_char_p, c_size_t]),
    "fl_device_cus": (c_int, []),
    "fl_eval_workspace_size": (c_size_t, [c_int64, c_int64, c_int, c_int]),
    "fl_eval_sym_workspace_size": (c_size_t, [c_int64, c_int, c_int]),
    "fl_eval": (c_int, [POINTER(fl_kernel), c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                        c_void_p, c_double, c_double, c_void_p, c_void_p, c_int, c_void_p, c_size_t, c_void_p]),
    "fl_eval_noalloc": (c_int, [POINTER(fl_kernel), c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                                c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_int, c_void_p]),
    "fl_bltc_weights": (c_int, [c_int64, c_void_p, c_void_p, c_int, c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                                c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fl_bltc_eval": (c_int, [POINTER(fl_kernel), c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                             c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_double, c_double,
                             c_double, c_void_p, c_void_p, c_void_p]),
    "fl_assemble": (c_int, [POINTER(fl_kernel), c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_double,
                            c_double, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p]),
    "fl_assemble_nearest": (c_int, [POINTER(fl_kernel), c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                    c_int64, c_void_p, c_void_p, c_double, c_void_p, c_int64, c_void_p, c_int64,
                                    c_void_p]),
    "fl_nearest": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fl_getrf_aux_size": (c_size_t, [c_int64, c_int]),
    "fl_getrf_workspace_size": (c_size_t, [c_int64, c_int]),
    "fl_sgetrf": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, POINTER(c_int), POINTER(c_int), c_void_p,
                          c_size_t, c_void_p]),
    "fl_dgetrf": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, POINTER(c_int), POINTER(c_int), c_void_p,
                          c_size_t, c_void_p]),
    "fl_sgetrf_async": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                c_void_p]),
    "fl_sgetri": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_void_p]),
    "fl_pad_solve": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "fl_refine_steps": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                                c_int64, c_int64, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "fl_pad_solve_group": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                   c_int64, c_void_p]),
    "fl_sgetrf_async_group": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                      c_void_p, c_size_t, c_void_p]),
    "fl_refine_steps_group": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                                      c_int64, c_int64, c_int, c_void_p, c_int64, c_int64, c_void_p, c_size_t,
                                      c_void_p]),
    "fl_refine_inv": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                              c_int64, c_double, c_int, POINTER(c_int), POINTER(c_double), c_void_p, c_size_t,
                              c_void_p]),
    "fl_sgetrf_ex": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, POINTER(c_int), POINTER(c_int), c_int,
                             c_void_p, c_size_t, c_void_p]),
    "fl_sgetrf_async_ex": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                   c_size_t, c_void_p]),
    "fl_ssytrf_async": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                                c_void_p]),
    "fl_dsytrf_async": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                c_void_p]),
    "fl_dgetrf_async": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                c_void_p]),
    "fl_getrs_workspace_size": (c_size_t, [c_int64, c_int64, c_int]),
    "fl_sgetrs": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_size_t,
                          c_void_p]),
    "fl_dgetrs": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_size_t,
                          c_void_p]),
    "fl_refine_workspace_size": (c_size_t, [c_int64, c_int64]),
    "fl_refine": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                          c_int64, c_double, c_int, POINTER(c_int), POINTER(c_double), c_void_p, c_size_t, c_void_p]),
    "fl_refine_frame": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, POINTER(c_double), c_void_p,
                                c_void_p, c_int64, c_int64, c_double, c_int, POINTER(c_int), POINTER(c_double),
                                c_void_p, c_size_t, c_void_p]),
    "fl_gmres_workspace_size": (c_size_t, [c_int64, c_int, c_int]),
    "fl_gmres": (c_int, [POINTER(fl_kernel), c_int64, c_void_p, c_void_p, c_int64, c_double, c_void_p, c_void_p,
                         c_double, c_int, c_int, c_int, POINTER(c_int), POINTER(c_double), c_void_p, c_size_t,
                         c_void_p]),
    "fl_gmres_bj_workspace_size": (c_size_t, [c_int64, c_int, c_int]),
    "fl_gmres_bj": (c_int, [POINTER(fl_kernel), c_int64, c_void_p, c_void_p, c_int64, c_double, c_void_p, c_void_p,
                            c_double, c_int, c_int, c_int, POINTER(c_int), POINTER(c_double), c_void_p, c_size_t,
                            c_void_p]),
    "fl_assemble_blocks": (c_int, [POINTER(fl_kernel), c_int64, c_int, c_void_p, c_void_p, c_int64, c_double,
                                   c_void_p, c_void_p]),
    "fl_block_lu": (c_int, [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fl_block_apply": (c_int, [c_int64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fl_sgemm": (c_int, [c_int64, c_int64, c_int64, c_float, c_void_p, c_int64, c_void_p, c_int64, c_float,
                         c_void_p, c_int64, c_void_p]),
    "fl_dgemm": (c_int, [c_int64, c_int64, c_int64, c_double, c_void_p, c_int64, c_void_p, c_int64, c_double,
                         c_void_p, c_int64, c_void_p]),
    "fl_swim_rhs": (c_int, [c_int64, c_void_p, POINTER(c_double), c_void_p, c_void_p, c_void_p]),
    "fl_swim_schur": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
}

_lib = None


def lib():
    """The loaded library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH.name} is not built; run `python -m fluidlets_amd.build` "
                              "(fluidlets_amd has no CPU fallback)")
        # torch first: the library then binds to the HIP runtime torch has
        # loaded instead of bringing a second copy into the process (two
        # runtimes cannot share the GPU)
        import torch  # noqa: F401
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGNATURES)


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    lib().fl_get_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(status: int, what: str) -> None:
    """Raise the Python exception the reference raises for this kind of failure."""
    if status == FL_OK:
        return
    msg = f"{what}: {last_error()}"
    if status == FL_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def kernel_struct(kernel) -> fl_kernel:
    """The device descriptor of a kernel object.

    In the reference a kernel is its Numba ``block`` function and the core
    loops call whatever ``block`` a subclass supplies
    (src/fluidlets/kernels/base.py:1-7, 14-19).  Here the pair math of each
    kernel is compiled into the HIP library, so only the three kernels it
    implements are accepted: Stokeslet3D, WallStokeslet3D and
    Brinkmanlet3D(alpha), or subclasses of them that keep their ``block``
    and ``kind``.  Any other kernel (a new ``Kernel`` subclass, or a
    subclass that supplies its own ``block``) raises NotImplementedError
    instead of silently computing a built-in kernel."""
    from .kernels import Brinkmanlet3D, Stokeslet3D, WallStokeslet3D

    for base in (WallStokeslet3D, Brinkmanlet3D, Stokeslet3D):
        if isinstance(kernel, base):
            cls = type(kernel)
            if cls.block is base.block and cls.kind == base.kind:
                alpha = float(getattr(kernel, "alpha", 0.0)) if base is Brinkmanlet3D else 0.0
                if not (alpha >= 0.0 and np.isfinite(alpha)):
                    raise ValueError("alpha must be real and non-negative")
                return fl_kernel(int(base.kind), 0, alpha)
            break
    raise NotImplementedError(
        f"{type(kernel).__name__} has no device implementation: fluidlets_amd runs the kernels compiled into "
        "libfluidlets_hip.so (Stokeslet3D, WallStokeslet3D, Brinkmanlet3D(alpha)); a kernel defined by its own "
        "`block` function cannot be dispatched to the GPU")
