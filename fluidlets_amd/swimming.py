"""Free swimming in Stokes flow on the device (drop-in for
src/fluidlets/swimming.py: SwimmerShape, SwimState, swimming_problem,
Trajectory, swim), plus a fixed-step midpoint RK2 integrator.

At every instant the unknowns are the forces f on the fluid at N force
points and the frame velocity w = [U; Omega]; with the nearest-neighbour
kernel matrix K (``assemble_nearest``) they solve (swimming.py:90-101)

    K f - U + (x - x0) x Omega = v,   sum count_n f_n = 0,   sum Xsum_n x f_n = 0.

K is assembled and factored on the GPU (FP64 LU, or FP32 LU with FP64
refinement); the seven right-hand sides [v | C] are solved with the factors
and the 6 x 6 Schur complement S = B K^{-1} C is formed and solved in FP64 by
the fl_swim_schur kernels.

Factor reuse across time steps ("mixed" precision, free-space kernels).  For
a kernel with G(R d) = R G(d) R^T (Stokeslet3D, Brinkmanlet3D) the lab-frame
matrix is the body-frame matrix rotated point by point, K = R K_body R^T,
so the FP32 factors of K_body depend only on the body shape, which a
periodic beat revisits.  They are kept in a FactorCache keyed by the body
shape; every step still assembles the exact lab-frame FP64 K and refines
against it (fl_refine_frame rotates each correction by the frame's
orthogonal polar factor), so the answer is the lab-frame solution to the
refinement tolerance whether the factors were reused or not.  When a shape
comes back a second time (its third visit), its factors are traded for the
explicit FP32 inverse (fl_sgetri, the same memory), so every later
correction is one pass over it (fl_refine_inv) instead of two chained
triangular sweeps.  Up to then the factors serve, since the inverse costs
40 % of a factorization more.  Concurrent swimmers sharing a cache make each
shape once (FactorCache.get_or_make), and swim_rk2 makes the factors of the
coming stage shapes on helper threads and streams ahead of the integrator,
since the body-frame shapes depend only on time.  A step's result depends
on the cache's history only within the refinement tolerance.
"""
from __future__ import annotations

import ctypes
import hashlib
import threading
from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import _device as D
from . import linalg
from ._lib import check, lib
from .cache import FactorCache
from .core.assemble import assemble_nearest_device, nearest_groups
from .kernels import Brinkmanlet3D, Stokeslet3D


def _cross_matrix(v: np.ndarray) -> np.ndarray:
    """(N, 3, 3) matrices C with C @ w = v x w for each row v (swimming.py:33-40)."""
    v = np.atleast_2d(v)
    C = np.zeros((len(v), 3, 3))
    C[:, 0, 1], C[:, 0, 2] = -v[:, 2], v[:, 1]
    C[:, 1, 0], C[:, 1, 2] = v[:, 2], -v[:, 0]
    C[:, 2, 0], C[:, 2, 1] = -v[:, 1], v[:, 0]
    return C


def _frame_rotation(B: np.ndarray) -> np.ndarray:
    """Orthogonal polar factor of the frame matrix (B itself when B is a
    rotation; the integrators do not re-orthonormalise B)."""
    u, _, vt = np.linalg.svd(B)
    return u @ vt


def _body_key(kernel, shape, nearest, eps, mu) -> str:
    """Cache key of a body-frame matrix: the kernel, mu, eps, the nearest map
    and the body shape on a grid of 2^-30 of its extent (shapes that differ
    only by rounding, e.g. one beat period later, share their factors)."""
    h = hashlib.sha1()
    h.update(f"{kernel!r}|{float(mu)!r}".encode())
    x = np.asarray(shape.x, np.float64)
    X = np.asarray(shape.X, np.float64)
    ext = max(float(np.abs(x).max(initial=0.0)), float(np.abs(X).max(initial=0.0)), 1e-300)
    q = 2.0 ** 30 / ext
    for a in (x, X):
        h.update(np.ascontiguousarray(np.rint(a * q), dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(nearest, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(np.asarray(eps, np.float64)).tobytes())
    return h.hexdigest()


def _rotation_invariant(kernel) -> bool:
    return isinstance(kernel, (Stokeslet3D, Brinkmanlet3D))


@dataclass
class SwimmerShape:
    """Swimmer geometry at one instant, in the body frame.

    x : (N, 3) force (collocation) points
    v : (N, 3) their body-frame velocities (the beat)
    X : (Q, 3) quadrature points
    """
    x: np.ndarray
    v: np.ndarray
    X: np.ndarray


@dataclass
class SwimState:
    U: np.ndarray            # frame velocity (lab)
    Omega: np.ndarray        # angular velocity (lab)
    forces: np.ndarray       # (N, 3) force on the fluid at each force point (lab)
    x: np.ndarray            # (N, 3) force points (lab)
    X: np.ndarray            # (Q, 3) quadrature points (lab)
    nearest: np.ndarray      # (Q,) nearest force point of each quadrature point


def _eps_points(eps, Q: int) -> np.ndarray:
    """Blob widths of the Q quadrature points: one value when all are equal."""
    eps_arr = np.broadcast_to(np.asarray(eps, float), (Q,))
    return eps_arr if not (Q > 1 and np.all(eps_arr == eps_arr[0])) else eps_arr[:1]


def _make_body_factors(kernel, shape: SwimmerShape, nearest, e, mu: float, stream, lu_hint):
    """FP32 factors of the body-frame matrix of `shape` (the factor cache's
    item for its key): assembled and factored on `stream` (torch's current
    stream carries the uploads), complete when this returns."""
    t = D.torch()
    dev = D.device()
    N = len(shape.x)
    xb, Xb = D.soa(np.asarray(shape.x, float), dev), D.soa(np.asarray(shape.X, float), dev)
    order, start = nearest_groups(nearest, N)
    order_t, start_t = t.from_numpy(order).to(dev), t.from_numpy(start).to(dev)
    eps_t, stride = D.eps_arg(e, len(shape.X), dev)

    def body32():
        return assemble_nearest_device(kernel, xb, Xb, eps_t, stride, order_t, start_t, N, mu=mu, fp32=True,
                                       fp64=False, stream=stream)[1]

    F = linalg.getrf(body32(), stream=stream, refill=lambda A: A.copy_(body32()), hint=lu_hint,
                     products="bf16x3")        # returns with the factors complete
    if F.info > 0:
        raise np.linalg.LinAlgError("Singular matrix")
    return (F, None), F.nbytes()


def swimming_problem(shape: SwimmerShape, x0, B, eps, mu=1.0, kernel=None, nearest=None,
                     precision: str = "fp64", stream=None, factor_cache: FactorCache | None = None,
                     lu_hint: dict | None = None) -> SwimState:
    """Solve the instantaneous free-swimming problem (reference signature).

    ``precision`` "fp64" (default) factors K in FP64; "mixed" uses FP32 MFMA
    factors with FP64 iterative refinement of all seven right-hand sides.
    With ``factor_cache`` ("mixed", Stokeslet3D or Brinkmanlet3D) the FP32
    factors of the body-frame matrix are taken from, or added to, the cache.
    ``lu_hint`` is shared by the solves of one trajectory (linalg.getrf): a
    swimmer whose matrix needs row interchanges tries the asynchronous LU
    once."""
    t = D.torch()
    kernel = Stokeslet3D() if kernel is None else kernel
    if np.iscomplexobj(mu):
        raise ValueError("swimming_problem needs a real viscosity")
    mu = float(mu)
    x0 = np.asarray(x0, float)
    B = np.asarray(B, float)
    x = x0 + np.asarray(shape.x, float) @ B.T
    X = x0 + np.asarray(shape.X, float) @ B.T
    v = np.asarray(shape.v, float) @ B.T
    dev = D.device()
    x_soa, X_soa = D.soa(x, dev), D.soa(X, dev)
    N, Q = len(x), len(X)
    if nearest is None:
        from ._nearest import nearest_device
        nearest = nearest_device(X_soa, x_soa, stream=stream).cpu().numpy()
    nearest = np.ascontiguousarray(nearest, dtype=np.int64)
    order, start = nearest_groups(nearest, N)
    e = _eps_points(eps, Q)
    eps_t, stride = D.eps_arg(e, Q, dev)
    order_t, start_t = t.from_numpy(order).to(dev), t.from_numpy(start).to(dev)
    body = precision == "mixed" and factor_cache is not None and _rotation_invariant(kernel)
    K64, K32 = assemble_nearest_device(kernel, x_soa, X_soa, eps_t, stride, order_t, start_t, N, mu=mu,
                                       fp32=(precision == "mixed" and not body), stream=stream)
    L = lib()
    # every device buffer stays referenced until its kernels are enqueued
    # (a temporary freed inside the argument list could be handed to the next
    # allocation of the same call)
    R = t.empty((3 * N, 7), dtype=t.float64, device=dev)
    v_soa = D.soa(v, dev)
    x0c = (ctypes.c_double * 3)(*x0.tolist())
    check(L.fl_swim_rhs(N, D.ptr(x_soa), x0c, D.ptr(v_soa), D.ptr(R), D.stream_handle(stream)), "fl_swim_rhs")
    if precision == "fp64":
        F = linalg.getrf(K64, stream=stream, hint=lu_hint,
                         refill=lambda A: assemble_nearest_device(kernel, x_soa, X_soa, eps_t, stride, order_t, start_t,
                                                                  N, mu=mu, K64=A, stream=stream))
        if F.info > 0:
            raise np.linalg.LinAlgError("Singular matrix")
        Y = linalg.getrs(F, R, stream=stream)
    elif precision == "mixed" and body:
        key = _body_key(kernel, shape, nearest, e, mu)
        item = factor_cache.get_or_make(
            key, lambda: _make_body_factors(kernel, shape, nearest, e, mu, stream, lu_hint))
        if item[1] is None and factor_cache.key_uses.get(key, 0) >= 3:
            # the shape keeps coming back (third visit): trade its factors for
            # the explicit inverse, so this and later visits correct with one
            # pass over it instead of two chained triangular sweeps
            F0 = item[0]

            def make_inverse():
                X_inv = linalg.getri(F0, stream=stream)
                (stream if stream is not None else t.cuda.current_stream()).synchronize()   # complete before shared
                return (None, X_inv), int(X_inv.numel() * X_inv.element_size())

            item = factor_cache.upgrade(key, item, make_inverse)
        F, Xinv = item[0], item[1]
        Y, _, _ = linalg.refine(K64, None if Xinv is not None else F, R, tol=1e-13, max_iter=30, stream=stream,
                                frame=_frame_rotation(B), inverse=Xinv)
    elif precision == "mixed":
        F = linalg.getrf(K32, stream=stream, refill=lambda A: A.copy_(K64), hint=lu_hint, products="bf16x3")
        if F.info > 0:
            raise np.linalg.LinAlgError("Singular matrix")
        Y, _, _ = linalg.refine(K64, F, R, tol=1e-13, max_iter=30, stream=stream)
    else:
        raise ValueError(f"unknown precision {precision!r}")
    count = np.bincount(nearest, minlength=N).astype(float)
    # per force point, the sum of its quadrature points (bincount adds in the
    # quadrature order, as np.add.at does, and is several times faster)
    Xsum = np.stack([np.bincount(nearest, weights=X[:, a], minlength=N) for a in range(3)], axis=1)
    f = t.empty((3 * N,), dtype=t.float64, device=dev)
    UW = t.empty((6,), dtype=t.float64, device=dev)
    S = t.empty((36,), dtype=t.float64, device=dev)
    count_d = t.from_numpy(count).to(dev)
    Xsum_d = D.soa(Xsum, dev)
    check(L.fl_swim_schur(N, D.ptr(Y), D.ptr(count_d), D.ptr(Xsum_d), D.ptr(f), D.ptr(UW), D.ptr(S),
                          D.stream_handle(stream)), "fl_swim_schur")
    w = UW.cpu().numpy()
    return SwimState(w[:3].copy(), w[3:].copy(), f.cpu().numpy().reshape(N, 3), x, X, nearest)


@dataclass
class Trajectory:
    t: np.ndarray            # (T,)
    x0: np.ndarray           # (T, 3) frame origin
    B: np.ndarray            # (T, 3, 3) frame axes as columns


def _frame_rhs(shape_at, eps, mu, kernel, nearest, precision):
    lu_hint: dict = {}

    def rhs(t, z):
        b1, b2 = z[3:6], z[6:9]
        B = np.c_[b1, b2, np.cross(b1, b2)]
        st = swimming_problem(shape_at(t), z[:3], B, eps, mu, kernel, nearest, precision=precision, lu_hint=lu_hint)
        return np.concatenate([st.U, np.cross(st.Omega, b1), np.cross(st.Omega, b2)])
    return rhs


def swim(shape_at: Callable[[float], SwimmerShape], t_span, eps, mu=1.0, x0=(0, 0, 0), B0=None, kernel=None,
         t_eval=None, rtol=1e-6, atol=1e-9, fixed_nearest=False, precision: str = "fp64") -> Trajectory:
    """Integrate a free swimmer's position and orientation (reference
    signature, swimming.py:112-148): scipy's RK45 on the host, every
    right-hand side a device swimming_problem."""
    from scipy.integrate import solve_ivp

    B0 = np.eye(3) if B0 is None else np.asarray(B0, float)
    nearest = None
    if fixed_nearest:
        from ._nearest import nearest_neighbors
        s0 = shape_at(t_span[0])
        nearest = nearest_neighbors(s0.X, s0.x)
    rhs = _frame_rhs(shape_at, eps, mu, kernel, nearest, precision)
    z0 = np.concatenate([np.asarray(x0, float), B0[:, 0], B0[:, 1]])
    sol = solve_ivp(rhs, t_span, z0, t_eval=t_eval, rtol=rtol, atol=atol)
    if not sol.success:
        raise RuntimeError(sol.message)
    b1, b2 = sol.y[3:6].T, sol.y[6:9].T
    Bt = np.stack([b1, b2, np.cross(b1, b2)], axis=2)
    return Trajectory(sol.t, sol.y[:3].T, Bt)


PREFETCH_DEPTH = 4      # stage shapes made ahead of the integrator
PREFETCH_WORKERS = 3    # helpers, each on its own stream, taking every third stage


def _prefetch_body_factors(shape_at, stage_times, kernel, nearest, e, mu, cache: FactorCache, progress, cond,
                           first: int = 0, step: int = 1):
    """swim_rk2's helper: make the body-frame factors of the stage shapes in
    order, at most PREFETCH_DEPTH ahead of the integrator, on a stream of its
    own.  The cache makes each shape once (FactorCache.get_or_make), so the
    integrator either finds a shape ready or waits for it; a failure here
    only ends the helper (the integrator then makes what it needs).  After
    2 PREFETCH_DEPTH shapes in a row that were already made (a beat already
    in the cache) the helper stops, leaving the interpreter to the
    integrator."""
    t = D.torch()
    stream = D.lease_stream()
    hint: dict = {}
    idle = 0
    try:
        for k in range(first, len(stage_times), step):
            tk = stage_times[k]
            if idle >= 2 * PREFETCH_DEPTH // step:
                return
            with cond:
                while not progress["stop"] and k >= progress["done"] + PREFETCH_DEPTH:
                    cond.wait()
                if progress["stop"]:
                    return
            shape = shape_at(tk)
            key = _body_key(kernel, shape, nearest, e, mu)
            made: list = []
            with t.cuda.stream(stream):
                cache.get_or_make(key, lambda: _make_body_factors(kernel, shape, nearest, e, mu, stream, hint),
                                  made, count=False)
            idle = 0 if made[0] else idle + 1
    except Exception:
        return
    finally:
        stream.synchronize()
        D.release_stream(stream)


def reuse_cache_bytes(n_points: int) -> int:
    """Bound of a swimmer's own factor cache: 128 body shapes' FP32 inverses
    (two solves per step, so a beat of 64 steps or fewer fits), at most 16 GiB."""
    n = 3 * int(n_points)
    return int(min(16 << 30, 128 * (4 * n * n + (1 << 20))))


def swim_rk2(shape_at: Callable[[float], SwimmerShape], t0: float, dt: float, steps: int, eps, mu=1.0,
             x0=(0, 0, 0), B0=None, kernel=None, fixed_nearest=True, precision: str = "fp64",
             stream=None, factor_cache: FactorCache | None = None, reuse_factors: bool = True,
             prefetch: bool = True) -> Trajectory:
    """Fixed-step explicit midpoint (RK2) integration of the frame:
    z_{n+1} = z_n + dt F(t_n + dt/2, z_n + dt/2 F(t_n, z_n)),  two device
    swimming_problem solves per step.

    With precision "mixed" and a free-space kernel the body-frame FP32
    factors are reused whenever the beat revisits a shape (``factor_cache``,
    or a cache of this call bounded by reuse_cache_bytes); reuse_factors=False
    factors every step.  The body-frame shapes depend only on time, so
    PREFETCH_WORKERS helper threads make the factors of the next
    PREFETCH_DEPTH stage shapes, each on a stream of its own, while the
    integrator solves (the factorizations of one swimmer are latency-bound
    and leave most of the GPU idle); they are joined before this returns.  prefetch=False leaves all making to the
    integrator."""
    B0 = np.eye(3) if B0 is None else np.asarray(B0, float)
    nearest = None
    s0 = shape_at(t0)
    if fixed_nearest:
        from ._nearest import nearest_neighbors
        nearest = nearest_neighbors(s0.X, s0.x)
    if not reuse_factors or precision != "mixed":
        factor_cache = None
    elif factor_cache is None:
        factor_cache = FactorCache(reuse_cache_bytes(len(s0.x)))
    lu_hint: dict = {}
    kern = Stokeslet3D() if kernel is None else kernel
    progress = {"done": 0, "stop": False}
    cond = threading.Condition()

    def F(t, z):
        b1, b2 = z[3:6], z[6:9]
        B = np.c_[b1, b2, np.cross(b1, b2)]
        st = swimming_problem(shape_at(t), z[:3], B, eps, mu, kernel, nearest, precision=precision, stream=stream,
                              factor_cache=factor_cache, lu_hint=lu_hint)
        with cond:
            progress["done"] += 1
            cond.notify_all()
        return np.concatenate([st.U, np.cross(st.Omega, b1), np.cross(st.Omega, b2)])

    helpers: list = []
    if (prefetch and factor_cache is not None and nearest is not None and _rotation_invariant(kern)
            and not np.iscomplexobj(mu)):
        stage_times = []
        t = t0
        for _ in range(steps):       # the integrator's own time arithmetic, so the shapes match
            stage_times += [t, t + 0.5 * dt]
            t += dt
        helpers = [threading.Thread(target=_prefetch_body_factors,
                                    args=(shape_at, stage_times, kern, nearest, _eps_points(eps, len(s0.X)),
                                          float(mu), factor_cache, progress, cond, i, PREFETCH_WORKERS))
                   for i in range(PREFETCH_WORKERS)]
        for h in helpers:
            h.start()
    z = np.concatenate([np.asarray(x0, float), B0[:, 0], B0[:, 1]])
    ts, zs = [t0], [z.copy()]
    t = t0
    try:
        for _ in range(steps):
            k1 = F(t, z)
            z = z + dt * F(t + 0.5 * dt, z + 0.5 * dt * k1)
            t += dt
            ts.append(t)
            zs.append(z.copy())
    finally:
        if helpers:
            with cond:
                progress["stop"] = True
                cond.notify_all()
            for h in helpers:
                h.join()
    zs = np.array(zs)
    b1, b2 = zs[:, 3:6], zs[:, 6:9]
    return Trajectory(np.array(ts), zs[:, :3], np.stack([b1, b2, np.cross(b1, b2)], axis=2))


class SwimmerSession:
    """A free swimmer advanced by one explicit midpoint (RK2) step per call,
    its state (time, frame origin, frame axes) kept between calls, for
    callers that interleave a trajectory's steps with other work (a serving
    loop's swimmer time-step requests).  ``step`` is one step of swim_rk2
    (the same time arithmetic and right-hand sides, without the helper
    threads): with precision "mixed" and a free-space kernel the body-frame
    factors and inverses live in ``factor_cache`` (shared between sessions,
    so a shape that any session has met costs refinement only), which the
    device keeps resident between the calls.  Steps of one session must run
    one at a time, in order."""

    def __init__(self, shape_at: Callable[[float], SwimmerShape], dt: float, eps, mu=1.0, x0=(0, 0, 0), B0=None,
                 kernel=None, t0: float = 0.0, precision: str = "mixed"):
        from ._nearest import nearest_neighbors
        self.shape_at, self.dt, self.eps, self.mu = shape_at, float(dt), eps, mu
        self.kernel, self.precision = kernel, precision
        B0 = np.eye(3) if B0 is None else np.asarray(B0, float)
        s0 = shape_at(t0)
        self.nearest = nearest_neighbors(s0.X, s0.x)
        self.t = float(t0)
        self.z = np.concatenate([np.asarray(x0, float), B0[:, 0], B0[:, 1]])
        self.steps = 0
        self.lu_hint: dict = {}

    def _F(self, t, z, stream, factor_cache):
        b1, b2 = z[3:6], z[6:9]
        B = np.c_[b1, b2, np.cross(b1, b2)]
        cache = factor_cache if self.precision == "mixed" else None
        st = swimming_problem(self.shape_at(t), z[:3], B, self.eps, self.mu, self.kernel, self.nearest,
                              precision=self.precision, stream=stream, factor_cache=cache, lu_hint=self.lu_hint)
        return np.concatenate([st.U, np.cross(st.Omega, b1), np.cross(st.Omega, b2)])

    def step(self, stream=None, factor_cache: FactorCache | None = None) -> np.ndarray:
        """Advance one step; returns [t, x0 (3), B (3 x 3, row-major)] after it."""
        dt, t, z = self.dt, self.t, self.z
        k1 = self._F(t, z, stream, factor_cache)
        z = z + dt * self._F(t + 0.5 * dt, z + 0.5 * dt * k1, stream, factor_cache)
        self.t, self.z = t + dt, z
        self.steps += 1
        b1, b2 = z[3:6], z[6:9]
        return np.concatenate([[self.t], z[:3], np.c_[b1, b2, np.cross(b1, b2)].ravel()])


def biflagellate(n_body: int = 1000, n_flag: int = 400, amplitude: float = 0.2, wavelength: float = 1.0,
                 length: float = 2.0, period: float = 1.0):
    """shape_at(t) of a model biflagellate: a 2:1 prolate spheroid body
    (semi-axes 0.5, 0.5, 1 along z, Fibonacci points) and two flagella of
    n_flag points leaving the anterior pole (z = 1) at +-30 degrees from the
    axis, each carrying a helical travelling wave of the given amplitude and
    wavelength (period T).  Force and quadrature points coincide."""
    from .geometry import spheroid

    body = spheroid(n_body, 0.5, 1.0)
    s = (np.arange(n_flag) + 1) * length / n_flag
    k = 2 * np.pi / wavelength
    om = 2 * np.pi / period

    def flagellum(t, side):
        ang = side * np.pi / 6
        d = np.array([np.sin(ang), 0.0, np.cos(ang)])
        e1 = np.array([np.cos(ang), 0.0, -np.sin(ang)])
        e2 = np.array([0.0, 1.0, 0.0])
        ph = k * s - om * t
        base = np.array([0.0, 0.0, 1.02])
        # amplitude ramps up from the base so the flagellum leaves the body smoothly
        amp = amplitude * (1 - np.exp(-(s / 0.2) ** 2))
        pts = base + s[:, None] * d + amp[:, None] * (np.cos(ph)[:, None] * e1 + side * np.sin(ph)[:, None] * e2)
        vel = amp[:, None] * om * (np.sin(ph)[:, None] * e1 - side * np.cos(ph)[:, None] * e2)
        return pts, vel

    def shape_at(t):
        p1, v1 = flagellum(t, +1)
        p2, v2 = flagellum(t, -1)
        x = np.vstack([body, p1, p2])
        v = np.vstack([np.zeros_like(body), v1, v2])
        return SwimmerShape(x=x, v=v, X=x)

    return shape_at
