"""Small resistance solves captured as HIP graphs.

A served resistance solve of at most GRAPH_MAX_POINTS points runs on a
padded system.  Its n x n matrix K (n = 3N) is the leading block of an
n_pad x n_pad buffer, n_pad = n rounded up to a multiple of BUCKET_ROWS (the
factorization runs fastest at multiples of 128), padded with the identity:
[[K, 0], [0, I]].  Partial pivoting never takes a
padding row in the first n columns (they are zero there) and the padding's
L21 and U12 blocks stay zero, so the factors are K's next to I.  The padded
right-hand sides are zero below row n, and so are the residuals, the
corrections and the solution, whose first n rows solve K x = b.

One graph per (bucket, stream) is captured (torch.cuda.CUDAGraph around the
library's launches, in thread-local capture mode so that other streams keep
working) and replayed for every later solve of that bucket on that stream.
A server captures every bucket for every slot stream when it is built, in
the constructing thread; a pool without streams captures on first use, and
only in the thread that created it (GraphPool).  It holds the asynchronous FP32 LU without interchanges
(fl_sgetrf_async_ex) and GRAPH_STEPS FP64 refinement corrections
(fl_refine_steps): some 40 to 90 kernel launches become one graph launch,
with one synchronization after it instead of one per correction.  The
assembly runs before the replay, outside the graph, because the kernel's
parameters are launch arguments.  The replay writes only the graph's own
buffers.  Up to GRAPH_MAX_POINTS every trailing update runs on the library's
own GEMM kernel, so the graph holds nothing but the library's kernels.

After the replay the host reads the LU's status word and the residuals.  A
matrix that needs row interchanges is factored again on the verified path
(fl_sgetrf_ex) and refined with the same fixed steps; a solve whose
residuals exceed the tolerance is refined again from the start with the
adaptive stopping rule (fl_refine).  Both run outside the graph.  The direct
launches (``replay=False``) are the graph's kernels with the graph's
arguments, and resident padded factors are refined by the same steps, so a
request's result is bit for bit the same whichever stream, graph or path
served it.

No counterpart in the reference, whose solves are scipy.linalg.solve calls
(src/fluidlets/core/solve.py:66).
"""
from __future__ import annotations

import threading

import numpy as np

from . import _device as D
from . import linalg
from ._lib import FL_FP32C, FL_LU_FP32, check, kernel_struct, lib

GRAPH_MAX_POINTS = 640     # n_pad <= 1920
BUCKET_ROWS = 128
GRAPH_STEPS = 2            # corrections inside the graph (what 1e-10 takes on resistance matrices of these sizes)
NRHS = 6                   # the six rigid-body modes


def bucket_rows(n_points: int) -> int:
    """Padded order of an N-point solve."""
    return -(-3 * int(n_points) // BUCKET_ROWS) * BUCKET_ROWS


def graph_bytes(n_pad: int, nrhs: int = NRHS) -> int:
    """Device bytes one graph's buffers hold."""
    L = lib()
    return int(12 * n_pad * n_pad + L.fl_getrf_aux_size(n_pad, FL_FP32C) + L.fl_getrf_workspace_size(n_pad, FL_FP32C)
               + L.fl_refine_workspace_size(n_pad, nrhs) + 24 * n_pad * nrhs + 4 * n_pad + 32 * (n_pad // 3) + 1024)


def pool_bytes(streams: int) -> int:
    """Largest footprint of a pool serving `streams` streams: one graph per
    bucket and stream."""
    return streams * sum(graph_bytes(BUCKET_ROWS * k) for k in range(1, bucket_rows(GRAPH_MAX_POINTS) // BUCKET_ROWS + 1))


def _finish(K64, F, b, x, res, x_host, n, tol, stream):
    """After the fixed corrections (residuals res, x_host[:n] on the host):
    adaptive refinement from the start if any column's residual exceeds tol
    (or is NaN).  Returns x[:n] (host) and the residuals."""
    res = np.array(res, dtype=np.float64)
    if not np.all(res <= tol):
        xa, _, res = linalg.refine(K64, F, b, tol=tol, stream=stream)
        x.copy_(xa)
        x_host[:n].copy_(x[:n], non_blocking=True)
        stream.synchronize()
    return x_host[:n].numpy().copy(), np.asarray(res, dtype=np.float64)


class SolveGraph:
    """The buffers of one padded order on one stream and the graph that pads,
    factors and refines in them.

    Inputs travel in one pinned staging buffer and one copy: points (3 x N),
    blob widths (N, or one value broadcast), N as a double, right-hand sides
    (N x nrhs).  Outputs come back in one copy: the residuals, the LU's
    status word and x."""

    def __init__(self, n_pad: int, stream, steps: int = GRAPH_STEPS, nrhs: int = NRHS, capture: bool = True):
        t = D.torch()
        dev = D.device()
        L = lib()
        if not 0 < n_pad <= bucket_rows(GRAPH_MAX_POINTS):
            raise ValueError(f"padded order {n_pad} outside (0, {bucket_rows(GRAPH_MAX_POINTS)}]")
        if not 1 <= nrhs <= 8 or steps < 1:
            raise ValueError("need 1 <= nrhs <= 8 and at least one correction")
        self.n_pad, self.steps, self.nrhs, self.stream = n_pad, int(steps), int(nrhs), stream
        P = n_pad // 3
        self.off_eps, self.off_n, self.off_rhs = 3 * P, 4 * P, 4 * P + 8
        self.stage_host = t.empty((self.off_rhs + n_pad * nrhs,), dtype=t.float64, pin_memory=True)
        self.stage = t.zeros((self.off_rhs + n_pad * nrhs,), dtype=t.float64, device=dev)
        self.K64 = t.empty((n_pad, n_pad), dtype=t.float64, device=dev)
        self.K32 = t.empty((n_pad, n_pad), dtype=t.float32, device=dev)
        self.F = linalg.LUFactor(self.K32, t.empty((n_pad,), dtype=t.int32, device=dev),
                                 t.empty((int(L.fl_getrf_aux_size(n_pad, FL_FP32C)),), dtype=t.uint8, device=dev),
                                 n_pad, extra={"async": True, "products": "fp32"})
        self.ws_lu = t.empty((int(L.fl_getrf_workspace_size(n_pad, FL_FP32C)),), dtype=t.uint8, device=dev)
        self.b = t.zeros((n_pad, nrhs), dtype=t.float64, device=dev)
        # out = [residuals (8) | status word (int32 in slot 8) | unused | x (n_pad x nrhs)]
        self.out = t.zeros((16 + n_pad * nrhs,), dtype=t.float64, device=dev)
        self.resid = self.out[:8]
        self.status = self.out.view(t.int32)[16:17]
        self.x = self.out[16:].view(n_pad, nrhs)
        self.out_host = t.empty((16 + n_pad * nrhs,), dtype=t.float64, pin_memory=True)
        self.x_host = self.out_host[16:].view(n_pad, nrhs)
        self.ws_ref = t.empty((int(L.fl_refine_workspace_size(n_pad, nrhs)),), dtype=t.uint8, device=dev)
        self.replays = 0
        self.graph = None
        if capture:
            g = t.cuda.CUDAGraph()
            with t.cuda.stream(stream):
                g.capture_begin(capture_error_mode="thread_local")
                try:
                    self._launch()
                finally:
                    g.capture_end()
            self.graph = g

    def _launch(self):
        """The graph's work: identity padding and right-hand sides, the
        asynchronous LU of K32, fixed-step refinement of K64 x = b."""
        L = lib()
        n = self.n_pad
        sh = D.stream_handle(self.stream)
        base = D.ptr(self.stage)
        check(L.fl_pad_solve(D.ptr(self.K64), D.ptr(self.K32), n, base + 8 * self.off_n, base + 8 * self.off_rhs,
                             D.ptr(self.b), self.nrhs, sh), "pad_solve")
        check(L.fl_sgetrf_async_ex(D.ptr(self.K32), n, n, D.ptr(self.F.ipiv), D.ptr(self.F.aux), D.ptr(self.status),
                                   FL_LU_FP32, D.ptr(self.ws_lu), self.ws_lu.numel(), sh), "getrf_async")
        linalg.refine_steps(self.K64, self.F, self.b, self.x, self.resid, self.steps, self.ws_ref, stream=self.stream)

    def solve(self, kernel, points, eps, rhs, tol: float = 1e-10, replay: bool = True):
        """x with K x = rhs, K the resistance matrix of ``kernel`` on the
        (N, 3) ``points`` with blob widths ``eps`` (mu = 1), rhs (3N, nrhs)
        on the host.  Returns (x (3N, nrhs) host, per-column relative
        residuals, the factors used: the graph's own or, after interchanges,
        new ones)."""
        t = D.torch()
        points = np.ascontiguousarray(points, dtype=np.float64)
        N = len(points)
        n = 3 * N
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if N == 0 or n > self.n_pad or rhs.shape != (n, self.nrhs):
            raise ValueError(f"{N} points and rhs {rhs.shape} do not fit padded order {self.n_pad}, {self.nrhs} columns")
        st = self.stream
        q = self.nrhs
        h = self.stage_host.numpy()        # free: the previous solve on this stream ended with a synchronization
        h[:n] = points.T.ravel()
        e = np.asarray(eps, dtype=np.float64)
        if e.size == 1:
            h[self.off_eps] = float(e.reshape(-1)[0])
            stride = 0
        else:
            h[self.off_eps:self.off_eps + N] = np.broadcast_to(e, (N,))
            stride = 1
        h[self.off_n] = float(n)
        used = self.off_rhs + n * q
        h[self.off_rhs:used] = rhs.ravel()
        back = 16 + n * q
        with t.cuda.stream(st):
            self.stage[:used].copy_(self.stage_host[:used], non_blocking=True)
            base = D.ptr(self.stage)
            check(lib().fl_assemble(kernel_struct(kernel), N, N, base, base, base + 8 * self.off_eps, stride, 1.0,
                                    0.0, D.ptr(self.K64), self.n_pad, D.ptr(self.K32), self.n_pad, None,
                                    D.stream_handle(st)), "fl_assemble")
            if replay and self.graph is not None:
                self.graph.replay()
                self.replays += 1
            else:
                self._launch()
            self.out_host[:back].copy_(self.out[:back], non_blocking=True)
            st.synchronize()
            F = self.F
            if int(self.out_host.view(t.int32)[16]) != 0:   # interchanges needed: the verified factorization
                self.K32.copy_(self.K64)
                F = linalg.getrf(self.K32, stream=st)
                if F.info > 0:
                    raise np.linalg.LinAlgError("Matrix is singular.")
                linalg.refine_steps(self.K64, F, self.b, self.x, self.resid, self.steps, self.ws_ref, stream=st)
                self.out_host[:back].copy_(self.out[:back], non_blocking=True)
                st.synchronize()
            return (*_finish(self.K64, F, self.b, self.x, self.out_host[:q].numpy(), self.x_host, n, tol, st), F)

    def resident(self, F):
        """Copies of the padded FP64 matrix and the factors F (this graph's or
        the verified path's), for a cache of resident factors."""
        K64 = self.K64.clone()
        G = linalg.LUFactor(F.lu.clone(), F.ipiv.clone(), F.aux.clone(), F.n, F.info, F.general_panels,
                            dict(F.extra))
        return K64, G


def solve_resident(K64, F, rhs, steps: int = GRAPH_STEPS, tol: float = 1e-10, stream=None):
    """The fixed-step solve of a graph on resident padded factors (K64 and F
    from ``SolveGraph.resident``): the same corrections on the same values,
    so the result is bit for bit the graph's."""
    t = D.torch()
    st = stream if stream is not None else t.cuda.current_stream()
    n_pad = int(K64.shape[0])
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    n, nrhs = rhs.shape
    if n > n_pad:
        raise ValueError("right-hand side longer than the padded order")
    with t.cuda.stream(st):
        dev = K64.device
        b = t.zeros((n_pad, nrhs), dtype=t.float64, device=dev)
        b[:n].copy_(t.from_numpy(rhs))
        x = t.empty((n_pad, nrhs), dtype=t.float64, device=dev)
        resid = t.zeros((8,), dtype=t.float64, device=dev)
        ws = t.empty((int(lib().fl_refine_workspace_size(n_pad, nrhs)),), dtype=t.uint8, device=dev)
        linalg.refine_steps(K64, F, b, x, resid, steps, ws, stream=st)
        res = resid[:nrhs].cpu().numpy()
        x_host = t.empty((n_pad, nrhs), dtype=t.float64)
        x_host[:n].copy_(x[:n])
        return _finish(K64, F, b, x, res, x_host, n, tol, st)


class GraphPool:
    """Captured small-solve graphs, one per (padded order, stream).

    With ``streams`` every bucket's graph is captured for every stream at
    construction, by the constructing thread, and ``get`` never captures:
    a server builds its pool before its worker threads start, so no capture
    overlaps another thread's runtime calls or another capture (a capture
    made in a worker thread while other workers factored and captured ended
    the process in the HIP runtime).  Without ``streams`` graphs are
    captured on first use, for single-threaded callers.  A stream serves one
    request at a time, so a graph is never replayed by two threads at once."""

    def __init__(self, steps: int = GRAPH_STEPS, streams=None):
        self.steps = steps
        self._graphs: dict = {}
        self._lock = threading.Lock()
        self.captures = 0
        self.lazy = streams is None
        self._owner = threading.get_ident()
        for s in streams or ():
            for k in range(1, bucket_rows(GRAPH_MAX_POINTS) // BUCKET_ROWS + 1):
                rows = BUCKET_ROWS * k
                self._graphs[(rows, D.stream_handle(s))] = SolveGraph(rows, s, self.steps)
                self.captures += 1
        if streams is not None:
            D.torch().cuda.synchronize()

    def get(self, n_points: int, stream) -> SolveGraph:
        if not 0 < n_points <= GRAPH_MAX_POINTS:
            raise ValueError(f"graphs serve 1 to {GRAPH_MAX_POINTS} points, not {n_points}")
        key = (bucket_rows(n_points), D.stream_handle(stream))
        with self._lock:
            g = self._graphs.get(key)
        if g is None:
            if not self.lazy:
                raise KeyError(f"no graph captured for {key[0]} rows on this stream")
            if threading.get_ident() != self._owner:
                # a capture overlapping other threads' runtime calls can end the
                # process, so a lazy pool captures only in its creating thread
                raise RuntimeError("a GraphPool without streams captures only in the thread that created it; "
                                   "build it with streams= to capture every graph up front for other threads")
            g = SolveGraph(key[0], stream, self.steps)
            with self._lock:
                self._graphs[key] = g
                self.captures += 1
        return g

    @property
    def replays(self) -> int:
        with self._lock:
            return sum(g.replays for g in self._graphs.values())

    def nbytes(self) -> int:
        with self._lock:
            return sum(graph_bytes(k[0]) for k in self._graphs)
