"""CPU oracle for the Fluidlets hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and bench.py's
cpu_baseline leg may import this module, and only as the checker (or as the
timed CPU baseline).  The product package ``fluidlets_amd`` never imports it.

Two independent restatements of the reference live here:

* the C loops in ``fl_oracle.c`` (loaded through ctypes), which follow the
  reference's Numba loops and block functions form by form;
* NumPy per-pair forms that restate the reference's ``velocity`` methods
  (``kernels/stokeslet3d.py:59-64``, ``kernels/images.py:107-139``,
  ``kernels/brinkmanlet3d.py:124-141``), written as matrix algebra so the two
  check each other, as the reference's own tests do.

The solver oracles restate ``core/solve.py:30-84`` (dense LU through
scipy.linalg.solve, GMRES through scipy.sparse.linalg.gmres) and
``swimming.py:66-102`` (the bordered free-swimmer system through
numpy.linalg.solve).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from math import factorial
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"

KINDS = {"stokeslet": 0, "wall": 1, "brinkman": 2}

_lib = None


def build() -> None:
    """Compile fl_oracle.c with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        L.flo_apply.argtypes = [ctypes.c_int, ctypes.c_double, i64, i64, dp, dp, dp, dp, dp, dp, dp]
        L.flo_fill.argtypes = [ctypes.c_int, ctypes.c_double, i64, i64, dp, dp, dp, dp]
        L.flo_fill_nearest.argtypes = [ctypes.c_int, ctypes.c_double, i64, i64, i64, dp, dp,
                                       ctypes.POINTER(ctypes.c_int64), dp, dp]
        L.flo_block.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_double, dp]
        L.flo_image_block.argtypes = [dp, dp, ctypes.c_double, dp]
        L.flo_brink_f1.argtypes = [ctypes.c_double]
        L.flo_brink_f1.restype = ctypes.c_double
        L.flo_brink_f2.argtypes = [ctypes.c_double]
        L.flo_brink_f2.restype = ctypes.c_double
        L.flo_set_threads.argtypes = [ctypes.c_int]
        L.flo_max_threads.restype = ctypes.c_int
        L.flo_init()
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n: int) -> None:
    lib().flo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().flo_max_threads())


# ---------------------------------------------------------------- input normalisation
# core/_common.py:7-23

def points(a, dim: int = 3) -> np.ndarray:
    a = np.atleast_2d(np.asarray(a, dtype=np.float64))
    if a.shape[1] != dim:
        raise ValueError(f"expected points of shape (N, {dim}), got {a.shape}")
    return np.ascontiguousarray(a)


def blob_widths(eps, n: int) -> np.ndarray:
    eps = np.asarray(eps, dtype=np.float64)
    return np.ascontiguousarray(np.broadcast_to(eps, (n,)))


def field_dtype(mu, *arrays):
    return np.result_type(np.asarray(mu), np.float64, *[np.asarray(a) for a in arrays])


# ---------------------------------------------------------------- kernel sums (C loops)

def kind_of(kernel) -> tuple[int, float]:
    """(kind, alpha) from a name ('stokeslet', 'wall', 'brinkman:<alpha>') or an
    object with ``kind``/``alpha`` attributes."""
    if isinstance(kernel, str):
        name, _, a = kernel.partition(":")
        return KINDS[name], float(a or 0.0)
    return int(kernel.kind), float(getattr(kernel, "alpha", 0.0))


def evaluate(kernel, targets, sources, strengths, eps, mu=1.0) -> np.ndarray:
    """core/matvec.py:24-49 on the C restatement of _apply."""
    kind, alpha = kind_of(kernel)
    targets = points(targets)
    sources = points(sources)
    eps = blob_widths(eps, sources.shape[0])
    dtype = field_dtype(mu, strengths)
    f = np.ascontiguousarray(np.asarray(strengths, dtype=dtype).reshape(sources.shape[0], 3))
    M, N = targets.shape[0], sources.shape[0]
    out_re = np.zeros((M, 3))
    if np.iscomplexobj(f):
        fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
        out_im = np.zeros((M, 3))
    else:
        fr, fi, out_im = f, None, None
    lib().flo_apply(kind, alpha, M, N, _p(targets), _p(sources), _p(eps), _p(fr), _p(fi),
                    _p(out_re), _p(out_im))
    out = out_re.astype(dtype) if out_im is None else out_re + 1j * out_im
    return out / mu


def assemble(kernel, targets, sources, eps, mu=1.0) -> np.ndarray:
    """core/assemble.py:27-46 on the C restatement of _fill."""
    kind, alpha = kind_of(kernel)
    targets = points(targets)
    sources = points(sources)
    eps = blob_widths(eps, sources.shape[0])
    K = np.empty((3 * targets.shape[0], 3 * sources.shape[0]))
    lib().flo_fill(kind, alpha, targets.shape[0], sources.shape[0], _p(targets), _p(sources), _p(eps), _p(K))
    return K.astype(field_dtype(mu)) / mu


def assemble_nearest(kernel, targets, quad_points, nearest, n_force, eps, mu=1.0) -> np.ndarray:
    """core/assemble.py:68-101 on the C restatement of _fill_nearest."""
    kind, alpha = kind_of(kernel)
    targets = points(targets)
    quad = points(quad_points)
    eps = blob_widths(eps, quad.shape[0])
    nearest = np.ascontiguousarray(nearest, dtype=np.int64)
    K = np.zeros((3 * targets.shape[0], 3 * n_force))
    lib().flo_fill_nearest(kind, alpha, targets.shape[0], quad.shape[0], int(n_force), _p(targets),
                           _p(quad), nearest.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(eps), _p(K))
    return K.astype(field_dtype(mu)) / mu


def nearest_neighbors(quad_points, force_points) -> np.ndarray:
    """core/assemble.py:61-65 (cKDTree query)."""
    from scipy.spatial import cKDTree
    return cKDTree(np.asarray(force_points, float)).query(np.asarray(quad_points, float))[1]


def image_block(x, x0, eps) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=float)
    x0 = np.ascontiguousarray(x0, dtype=float)
    G = np.empty(9)
    lib().flo_image_block(_p(x), _p(x0), float(eps), _p(G))
    return G.reshape(3, 3)


def brink_f1(z: float) -> float:
    return float(lib().flo_brink_f1(float(z)))


def brink_f2(z: float) -> float:
    return float(lib().flo_brink_f2(float(z)))


# ---------------------------------------------------------------- NumPy per-pair forms

def stokeslet_velocity(x, x0, f, eps, mu=1.0):
    """kernels/stokeslet3d.py:59-64."""
    d = np.asarray(x, float) - np.asarray(x0, float)
    f = np.asarray(f)
    r2 = d @ d
    q = (r2 + eps**2) ** 1.5
    return ((r2 + 2 * eps**2) * f + (f @ d) * d) / (8 * np.pi * mu * q)


def image_velocity(x, x0, f, eps, mu=1.0):
    """kernels/images.py:107-139: the image system written as matrix algebra."""
    x = np.asarray(x, float)
    x0 = np.asarray(x0, float)
    f = np.asarray(f)
    h0 = x0[2]
    d = x - np.array([x0[0], x0[1], -h0])
    e2 = eps**2
    Q = d @ d + e2
    I = np.eye(3)
    D = np.outer(d, d)
    Z = np.diag([-1.0, -1.0, 1.0])
    ez = np.array([0.0, 0.0, 1.0])
    stokeslet = -((0.5 / Q**0.5 + 0.5 * e2 / Q**1.5) * I + 0.5 / Q**1.5 * D)
    dipole = (-h0**2 * (1 / Q**1.5 - 3 * e2 / Q**2.5) * I + 3 * h0**2 / Q**2.5 * D) @ Z
    Hdb1 = 2 * h0 * (-1.5 * e2 / Q**2.5 - 0.5 / Q**1.5)
    Hdb2 = h0 / Q**1.5
    Hdb3 = -3 * h0 / Q**2.5
    # doublet: rows 0-1 of the x,y columns carry -d3 (Hdb2 I + Hdb3 D); row 2 of
    # the x,y columns carries -d_j (Hdb1 + d3^2 Hdb3); the z column carries
    # d_i (Hdb2 + d3^2 Hdb3), plus d3 (Hdb1 + Hdb2) on the diagonal.
    top = -d[2] * (Hdb2 * I[:2, :2] + Hdb3 * D[:2, :2])
    bottom = -(Hdb1 + d[2] ** 2 * Hdb3) * d[:2]
    zcol = d * (Hdb2 + d[2] ** 2 * Hdb3) + np.array([0.0, 0.0, d[2] * (Hdb1 + Hdb2)])
    doublet = np.zeros((3, 3))
    doublet[:2, :2] = top
    doublet[2, :2] = bottom
    doublet[:, 2] = zcol
    Hr1 = 3 * h0 * e2 / Q**2.5
    rotlet = Hr1 * (d[2] * np.diag([1.0, 1.0, 0.0]) - np.outer(ez, [d[0], d[1], 0.0]))
    return (stokeslet + dipole + doublet + rotlet) @ f / (4 * np.pi * mu)


def wall_velocity(x, x0, f, eps, mu=1.0):
    """kernels/images.py:102-104."""
    return stokeslet_velocity(x, x0, f, eps, mu) + image_velocity(x, x0, f, eps, mu)


def brinkman_velocity(alpha, x, x0, f, eps, mu=1.0):
    """kernels/brinkmanlet3d.py:124-141 (closed forms only)."""
    d = np.asarray(x, float) - np.asarray(x0, float)
    f = np.asarray(f)
    R = np.sqrt(d @ d + eps**2)
    a = alpha
    if a == 0.0:
        h1, h2 = 1 / (8 * np.pi * R), 1 / (8 * np.pi * R**3)
    else:
        E = np.exp(-a * R)
        h1 = (E * (1 + 1 / (a * R) + 1 / (a * R) ** 2) - 1 / (a * R) ** 2) / (4 * np.pi * R)
        h2 = (3 / (a * R) ** 2 - E * (1 + 3 / (a * R) + 3 / (a * R) ** 2)) / (4 * np.pi * R**3)
    H1 = h1 + eps**2 * h2
    return (H1 * f + h2 * (f @ d) * d) / mu


def pairwise_sum(velocity, X, X0, F, eps):
    """Sum of per-pair reference velocities, the reference tests' check form."""
    eps = np.broadcast_to(np.asarray(eps, float), (len(X0),))
    return sum(np.array([velocity(x, X0[j], F[j], eps[j]) for x in X]) for j in range(len(X0)))


# ---------------------------------------------------------------- series coefficients
# kernels/brinkmanlet3d.py:43-49

N_TERMS = 24
C1 = np.array([(-1) ** n * (1 / factorial(n + 2) - 1 / factorial(n + 1) + 1 / factorial(n))
               for n in range(N_TERMS)])
C2 = np.array([(-1) ** (n + 1) * (3 / factorial(n + 2) - 3 / factorial(n + 1) + 1 / factorial(n))
               for n in range(N_TERMS)])


# ---------------------------------------------------------------- solvers

@dataclass
class SolveInfo:
    method: str
    iterations: int
    residual: float


def resistance(kernel, X, velocities, eps, mu=1.0, method="lu", rtol=1e-10, restart=60, maxiter=200):
    """core/solve.py:30-84 with the C kernel sums and LAPACK / scipy GMRES."""
    import scipy.linalg
    import scipy.sparse.linalg as spla

    X = points(X)
    n = X.shape[0]
    eps = blob_widths(eps, n)
    U = np.asarray(velocities).reshape(-1)
    dtype = field_dtype(mu, U)
    if method == "lu":
        K = assemble(kernel, X, X, eps, mu)
        f = scipy.linalg.solve(K, U.astype(dtype))
        iterations = 1
    elif method == "gmres":
        def matvec(v):
            return evaluate(kernel, X, X, v.reshape(n, -1), eps, mu).ravel()
        A = spla.LinearOperator((U.size, U.size), matvec=matvec, dtype=dtype)
        count = [0]
        f, flag = spla.gmres(A, U.astype(dtype), rtol=rtol, atol=0.0, restart=restart, maxiter=maxiter,
                             callback=lambda _: count.__setitem__(0, count[0] + 1), callback_type="pr_norm")
        if flag != 0:
            raise RuntimeError(f"GMRES did not converge (flag {flag})")
        iterations = count[0]
    else:
        raise ValueError(f"unknown method {method!r}")
    resid = np.linalg.norm(evaluate(kernel, X, X, f.reshape(n, -1), eps, mu).ravel() - U) / np.linalg.norm(U)
    return f.reshape(n, -1), SolveInfo(method, iterations, float(resid))


def cross_matrix(v):
    """swimming.py:33-40."""
    v = np.atleast_2d(v)
    C = np.zeros((len(v), 3, 3))
    C[:, 0, 1], C[:, 0, 2] = -v[:, 2], v[:, 1]
    C[:, 1, 0], C[:, 1, 2] = v[:, 2], -v[:, 0]
    C[:, 2, 0], C[:, 2, 1] = -v[:, 1], v[:, 0]
    return C


def swimming_problem(kernel, x_body, v_body, X_body, x0, B, eps, mu=1.0, nearest=None):
    """swimming.py:66-102: the bordered (3N+6) system solved densely in FP64.

    Returns (U, Omega, forces (N,3), nearest)."""
    x0 = np.asarray(x0, float)
    B = np.asarray(B, float)
    x = x0 + np.asarray(x_body, float) @ B.T
    X = x0 + np.asarray(X_body, float) @ B.T
    v = np.asarray(v_body, float) @ B.T
    if nearest is None:
        nearest = nearest_neighbors(X, x)
    N = len(x)
    A = np.zeros((3 * N + 6, 3 * N + 6))
    A[:3 * N, :3 * N] = assemble_nearest(kernel, x, X, nearest, N, eps, mu)
    A[:3 * N, 3 * N:3 * N + 3] = -np.tile(np.eye(3), (N, 1))
    A[:3 * N, 3 * N + 3:] = cross_matrix(x - x0).reshape(3 * N, 3)
    count = np.bincount(nearest, minlength=N).astype(float)
    Xsum = np.zeros((N, 3))
    np.add.at(Xsum, nearest, X)
    A[3 * N:3 * N + 3, :3 * N] = np.kron(count, np.eye(3))
    A[3 * N + 3:, :3 * N] = np.hstack(list(cross_matrix(Xsum)))
    rhs = np.concatenate([v.ravel(), np.zeros(6)])
    sol = np.linalg.solve(A, rhs)
    return sol[3 * N:3 * N + 3], sol[3 * N + 3:], sol[:3 * N].reshape(N, 3), nearest


def grand_resistance(forces_by_mode, X, center):
    """6x6 grand resistance matrix from six solved force fields (pattern of
    examples/translating_sphere/run.py:45-48): column m = [sum f; sum (x-c) x f]."""
    R = np.zeros((6, 6))
    r = np.asarray(X, float) - np.asarray(center, float)
    for m, f in enumerate(forces_by_mode):
        R[:3, m] = f.sum(axis=0)
        R[3:, m] = np.cross(r, f).sum(axis=0)
    return R


def rigid_mode_velocities(X, center):
    """Unit translations e_k and rotations e_k x (x - c), k = 0..2: (6, N, 3)."""
    X = np.asarray(X, float)
    r = X - np.asarray(center, float)
    out = np.zeros((6, len(X), 3))
    for k in range(3):
        out[k, :, k] = 1.0
        e = np.zeros(3)
        e[k] = 1.0
        out[3 + k] = np.cross(e, r)
    return out


if os.environ.get("FL_ORACLE_THREADS"):
    set_threads(int(os.environ["FL_ORACLE_THREADS"]))
