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

Grouped solves (GroupSolveGraph): a server with many small requests
waiting solves up to `members` of one bucket together.  Each member keeps
its own slab of buffers and every kernel is launched once for the group
(blockIdx.z selects the member), so a group of eight costs about what one
solve costs while the GPU is not full, and each member's result is bit for
bit its single solve's.

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


def pool_bytes(streams: int, members: int = 1) -> int:
    """Largest footprint of a pool serving `streams` streams: one graph (or
    group of `members` slabs) per bucket and stream."""
    per = group_bytes if members > 1 else (lambda rows, _: graph_bytes(rows))
    return streams * sum(per(BUCKET_ROWS * k, members) for k in range(1, bucket_rows(GRAPH_MAX_POINTS) // BUCKET_ROWS + 1))


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


def _align(n: int, a: int = 256) -> int:
    return -(-int(n) // a) * a


def _slab_layout(n_pad: int, nrhs: int) -> list:
    """(buffer, bytes) of one group member's slab, in slab order: staging
    [points 3 x P | eps P | N | rhs], K64, K32, pivots, factor aux, LU
    workspace, right-hand sides, outputs [residuals 8 | status | x],
    refinement workspace (P = n_pad / 3)."""
    L = lib()
    return [("stage", 8 * (4 * (n_pad // 3) + 8 + n_pad * nrhs)), ("K64", 8 * n_pad * n_pad),
            ("K32", 4 * n_pad * n_pad), ("ipiv", 4 * n_pad), ("aux", int(L.fl_getrf_aux_size(n_pad, FL_FP32C))),
            ("ws_lu", int(L.fl_getrf_workspace_size(n_pad, FL_FP32C))), ("b", 8 * n_pad * nrhs),
            ("out", 8 * (16 + n_pad * nrhs)), ("ws_ref", int(L.fl_refine_workspace_size(n_pad, nrhs)))]


class GroupSolveGraph:
    """Up to ``members`` solves of one padded order on one stream, every
    kernel of the graph launched once for the whole group.

    Each member owns one slab of a single device allocation (``slab_bytes``
    apart, every buffer at the same offset in its slab): staging, K64, K32,
    pivots, factor aux, workspaces, right-hand sides and outputs.  The
    grouped entry points (fl_pad_solve_group, fl_sgetrf_async_group,
    fl_refine_steps_group) run member z's problem in workgroups with
    blockIdx.z = z on the pointers of slab z, so each member's result is bit
    for bit what SolveGraph computes for it alone.  One graph is captured per
    group size 1..members; the assemblies (one launch per member, since their
    sizes are launch arguments) run before the replay, and one pitched copy
    each way moves the members' staging and outputs."""

    def __init__(self, n_pad: int, stream, members: int = 8, steps: int = GRAPH_STEPS, nrhs: int = NRHS,
                 capture: bool = True):
        t = D.torch()
        dev = D.device()
        if not 0 < n_pad <= bucket_rows(GRAPH_MAX_POINTS):
            raise ValueError(f"padded order {n_pad} outside (0, {bucket_rows(GRAPH_MAX_POINTS)}]")
        if not 1 <= nrhs <= 8 or steps < 1 or not 1 <= members <= 64:
            raise ValueError("need 1 <= nrhs <= 8, at least one correction and 1 to 64 members")
        self.n_pad, self.steps, self.nrhs, self.stream, self.members = n_pad, int(steps), int(nrhs), stream, members
        P = n_pad // 3
        self.off_eps, self.off_n, self.off_rhs = 3 * P, 4 * P, 4 * P + 8
        self.stage_len = self.off_rhs + n_pad * nrhs            # doubles
        self.out_len = 16 + n_pad * nrhs                        # doubles: residuals, status, x
        sizes = _slab_layout(n_pad, nrhs)
        self.off = {}
        o = 0
        for name, nbytes in sizes:
            self.off[name] = o
            o += _align(nbytes)
        self.size = dict(sizes)
        self.slab_bytes = o
        self.slab = t.zeros((members, o), dtype=t.uint8, device=dev)
        self.stage_host = t.zeros((members, self.stage_len), dtype=t.float64, pin_memory=True)
        self.out_host = t.zeros((members, self.out_len), dtype=t.float64, pin_memory=True)
        self.replays = 0
        self._members = [self._member_views(z) for z in range(members)]
        self.graphs: dict = {}
        if capture:
            for c in range(1, members + 1):
                g = t.cuda.CUDAGraph()
                with t.cuda.stream(stream):
                    g.capture_begin(capture_error_mode="thread_local")
                    try:
                        self._launch(c)
                    finally:
                        g.capture_end()
                self.graphs[c] = g

    def _view(self, z: int, name: str, dtype, shape):
        o = self.off[name]
        t = D.torch()
        itemsize = t.empty((), dtype=dtype).element_size()
        count = int(np.prod(shape))
        return self.slab[z, o:o + count * itemsize].view(dtype).view(*shape)

    def member(self, z: int) -> dict:
        """Member z's buffers as tensors (views into its slab)."""
        return self._members[z]

    def _member_views(self, z: int) -> dict:
        t = D.torch()
        n, q = self.n_pad, self.nrhs
        out = self._view(z, "out", t.float64, (self.out_len,))
        F = linalg.LUFactor(self._view(z, "K32", t.float32, (n, n)), self._view(z, "ipiv", t.int32, (n,)),
                            self._view(z, "aux", t.uint8, (self.size["aux"],)), n,
                            extra={"async": True, "products": "fp32"})
        return {"K64": self._view(z, "K64", t.float64, (n, n)), "F": F, "b": self._view(z, "b", t.float64, (n, q)),
                "out": out, "resid": out[:8], "x": out[16:].view(n, q),
                "ws_ref": self._view(z, "ws_ref", t.uint8, (self.size["ws_ref"],))}

    def _launch(self, count: int):
        """The group's work for its first `count` members: identity padding
        and right-hand sides, the asynchronous FP32 LU, fixed-step FP64
        refinement."""
        L = lib()
        n, q, S = self.n_pad, self.nrhs, self.slab_bytes
        sh = D.stream_handle(self.stream)
        base = D.ptr(self.slab)
        at = self.off
        check(L.fl_pad_solve_group(base + at["K64"], base + at["K32"], n, base + at["stage"] + 8 * self.off_n,
                                   base + at["stage"] + 8 * self.off_rhs, base + at["b"], q, count, S, sh),
              "pad_solve_group")
        check(L.fl_sgetrf_async_group(base + at["K32"], n, n, base + at["ipiv"], base + at["aux"],
                                      base + at["out"] + 64, count, S, base + at["ws_lu"], self.size["ws_lu"], sh),
              "getrf_async_group")
        check(L.fl_refine_steps_group(base + at["K64"], n, base + at["K32"], n, base + at["aux"], n, base + at["b"],
                                      base + at["out"] + 8 * 16, q, q, self.steps, base + at["out"], count, S,
                                      base + at["ws_ref"], self.size["ws_ref"], sh), "refine_steps_group")

    def solve_many(self, problems, tol: float = 1e-10, replay: bool = True, raise_errors: bool = True):
        """Solve each (kernel, points, eps, rhs) of `problems` (at most
        ``members``, each fitting the padded order; rhs (3N, nrhs) on the
        host) as SolveGraph.solve does.  Returns one (x, residuals, factors)
        per problem, in order; the factors are views into the member's slab
        (valid until the group's next solve) or, after interchanges, new
        ones.  With raise_errors=False a member whose solve fails after the
        group's launch (a singular matrix) gets its exception in place of its
        result and the other members still get theirs."""
        t = D.torch()
        m = len(problems)
        if not 1 <= m <= self.members:
            raise ValueError(f"a group solves 1 to {self.members} problems, not {m}")
        q, st = self.nrhs, self.stream
        h = self.stage_host.numpy()        # free: the previous solve on this stream ended with a synchronization
        dims = []
        for z, (kernel, points, eps, rhs) in enumerate(problems):
            points = np.ascontiguousarray(points, dtype=np.float64)
            N = len(points)
            n = 3 * N
            rhs = np.ascontiguousarray(rhs, dtype=np.float64)
            if N == 0 or n > self.n_pad or rhs.shape != (n, q):
                raise ValueError(f"{N} points and rhs {rhs.shape} do not fit padded order {self.n_pad}, {q} columns")
            h[z, :n] = points.T.ravel()
            e = np.asarray(eps, dtype=np.float64)
            if e.size == 1:
                h[z, self.off_eps] = float(e.reshape(-1)[0])
                stride = 0
            else:
                h[z, self.off_eps:self.off_eps + N] = np.broadcast_to(e, (N,))
                stride = 1
            h[z, self.off_n] = float(n)
            h[z, self.off_rhs:self.off_rhs + n * q] = rhs.ravel()
            dims.append((kernel, N, stride))
        S = self.slab_bytes
        base = D.ptr(self.slab)
        stage_d = base + self.off["stage"]
        with t.cuda.stream(st):
            D.copy_rows_async(stage_d, S, D.ptr(self.stage_host), 8 * self.stage_len, 8 * self.stage_len, m, "h2d", st)
            for z, (kernel, N, stride) in enumerate(dims):
                sz = stage_d + z * S
                check(lib().fl_assemble(kernel_struct(kernel), N, N, sz, sz, sz + 8 * self.off_eps, stride, 1.0, 0.0,
                                        base + z * S + self.off["K64"], self.n_pad, base + z * S + self.off["K32"],
                                        self.n_pad, None, D.stream_handle(st)), "fl_assemble")
            if replay and m in self.graphs:
                self.graphs[m].replay()
                self.replays += 1
            else:
                self._launch(m)
            D.copy_rows_async(D.ptr(self.out_host), 8 * self.out_len, base + self.off["out"], S, 8 * self.out_len, m,
                              "d2h", st)
            st.synchronize()
            results = []
            oh = self.out_host.numpy()
            status = oh.view(np.int32)[:, 16]
            for z, (kernel, N, _) in enumerate(dims):
                n = 3 * N
                res = oh[z, :q]
                if status[z] == 0 and np.all(res <= tol):   # the common case, read straight from the copy
                    results.append((oh[z, 16:16 + n * q].reshape(n, q).copy(), res.copy(), self._members[z]["F"]))
                    continue
                try:
                    results.append(self._finish_member(z, n, tol))
                except (np.linalg.LinAlgError, RuntimeError) as exc:
                    if raise_errors:
                        raise
                    results.append(exc)
            return results

    def _finish_member(self, z: int, n: int, tol: float):
        """Member z after the group's launch and read-back: the verified
        factorization if its LU needed interchanges, adaptive refinement if
        the fixed corrections fell short."""
        t = D.torch()
        st, q = self.stream, self.nrhs
        oh = self.out_host
        mb = self.member(z)
        F = mb["F"]
        if int(oh[z].view(t.int32)[16]) != 0:   # interchanges needed: the verified factorization
            F.lu.copy_(mb["K64"])
            F = linalg.getrf(F.lu, stream=st)
            if F.info > 0:
                raise np.linalg.LinAlgError("Matrix is singular.")
            linalg.refine_steps(mb["K64"], F, mb["b"], mb["x"], mb["resid"], self.steps, mb["ws_ref"], stream=st)
            oh[z].copy_(mb["out"], non_blocking=True)
            st.synchronize()
        x, res = _finish(mb["K64"], F, mb["b"], mb["x"], oh[z, :q].numpy(), oh[z, 16:].view(self.n_pad, q), n, tol, st)
        return x, res, F

    def resident(self, z: int, F):
        """Copies of member z's padded FP64 matrix and the factors F (its own
        or the verified path's), for a cache of resident factors."""
        K64 = self.member(z)["K64"].clone()
        G = linalg.LUFactor(F.lu.clone(), F.ipiv.clone(), F.aux.clone(), F.n, F.info, F.general_panels,
                            dict(F.extra))
        return K64, G


def group_bytes(n_pad: int, members: int, nrhs: int = NRHS) -> int:
    """Device bytes of a GroupSolveGraph's slabs."""
    return int(members * sum(_align(b) for _, b in _slab_layout(n_pad, nrhs)))


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

    def __init__(self, steps: int = GRAPH_STEPS, streams=None, members: int = 1):
        self.steps = steps
        self.members = int(members)
        self._graphs: dict = {}
        self._lock = threading.Lock()
        self.captures = 0
        self.lazy = streams is None
        self._owner = threading.get_ident()
        if self.members > 1 and streams is None:
            raise ValueError("grouped graphs are captured up front: give the pool its streams")
        for s in streams or ():
            for k in range(1, bucket_rows(GRAPH_MAX_POINTS) // BUCKET_ROWS + 1):
                rows = BUCKET_ROWS * k
                if self.members > 1:
                    g = GroupSolveGraph(rows, s, self.members, self.steps)
                    self.captures += len(g.graphs)
                else:
                    g = SolveGraph(rows, s, self.steps)
                    self.captures += 1
                self._graphs[(rows, D.stream_handle(s))] = g
        if streams is not None:
            D.torch().cuda.synchronize()

    def get_group(self, n_points: int, stream) -> GroupSolveGraph:
        """The grouped graphs of the bucket of `n_points` on `stream` (a pool
        built with members > 1)."""
        if self.members <= 1:
            raise ValueError("this pool holds single-solve graphs")
        if not 0 < n_points <= GRAPH_MAX_POINTS:
            raise ValueError(f"graphs serve 1 to {GRAPH_MAX_POINTS} points, not {n_points}")
        return self._graphs[(bucket_rows(n_points), D.stream_handle(stream))]

    def get(self, n_points: int, stream) -> SolveGraph:
        if self.members > 1:
            raise ValueError("this pool holds grouped graphs: use get_group")
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
            if self.members > 1:
                return sum(group_bytes(k[0], self.members) for k in self._graphs)
            return sum(graph_bytes(k[0]) for k in self._graphs)
