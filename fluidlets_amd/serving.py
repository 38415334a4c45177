"""Continuous-batching solve server on one GPU.

Requests (rigid-body resistance solves of any size, free-swimmer time
stepping) are admitted as they arrive whenever a slot is free and their HBM
footprint fits the budget; each slot is a long-lived worker thread driving
its own HIP stream (created for the server, default priority), so admitted
requests overlap on the device.  The
HIP library calls release the GIL, and every kernel is deterministic (fixed
summation orders, no floating-point atomics), so a request's result does not
depend on what else is in flight: tests check it bit for bit, alone and in a
mix.

Resident factors: resistance requests carry a geometry key; the assembled
FP64 matrix and the factors stay on the device in an LRU cache (counted in
the memory budget), so another right-hand side on the same geometry costs
only the triangular solves (and refinement).  Mixed-precision swimmers keep
the FP32 inverse of each body shape's matrix they meet (in that cache when the
server has one, else in a cache of their own, counted in their footprint), so
the steps of a periodic beat after the first period cost only the lab-frame
assembly, refinement with one pass over the inverse per correction, and the
6 x 6 Schur step.

Latency is measured from a request's arrival to its completion (queueing
included); the server reports p50/p99 and throughput.
"""
from __future__ import annotations

import hashlib
import heapq
import queue
import threading
import time
from contextlib import nullcontext
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _device as D
from .cache import FactorCache


@dataclass
class Request:
    kind: str                              # "resistance" | "swim_step" | "swimmer" (a whole trajectory)
    payload: dict
    arrival: float = 0.0                   # seconds, relative to the trace start
    rid: int = 0


@dataclass
class Result:
    rid: int
    kind: str
    value: Any
    arrival: float
    start: float
    end: float
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def footprint_bytes(req: Request, precision: str, shared_cache: bool = False) -> int:
    """HBM bytes a request holds while it runs: the FP64 matrix (8 n^2), the
    FP32 copy for the mixed path (4 n^2), plus O(n) vectors and workspaces;
    a mixed-precision swimmer also holds the FP32 inverse it makes next to
    its factors and the factors its helpers make ahead (4 n^2 each) and,
    without the server's shared cache, its own cache of body-frame
    factors."""
    if req.kind == "resistance":
        from .core.assemble import padded_order
        n = padded_order(3 * len(req.payload["points"]))
    elif req.kind == "swim_step":
        n = 3 * req.payload["spec"]["n_points"]
        per = 8 + (4 if precision == "mixed" else 8)
        extra = 4 * n * n if precision == "mixed" else 0     # an inverse made next to its factors
        if precision == "mixed" and not shared_cache:
            extra += 4 * n * n                                # no shared cache: the step's own factors
        return int(per * n * n + 512 * n + (64 << 20) + extra)
    else:
        n = 3 * req.payload["n_points"]
    per = 8 + (4 if precision == "mixed" else 8)
    extra = 0
    if req.kind == "swimmer" and precision == "mixed":
        from .swimming import PREFETCH_WORKERS
        extra += 4 * (1 + PREFETCH_WORKERS) * n * n   # its inverse next to its factors; its helpers' factors
    if req.kind == "swimmer" and precision == "mixed" and not shared_cache:
        from .swimming import reuse_cache_bytes
        extra += reuse_cache_bytes(req.payload["n_points"])   # the swimmer's own cache
    return int(per * n * n + 512 * n + (64 << 20) + extra)


def geometry_key(kernel, points, eps) -> str:
    h = hashlib.sha1()
    h.update(repr(kernel).encode())
    h.update(np.ascontiguousarray(points, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(np.asarray(eps, dtype=np.float64)).tobytes())
    return h.hexdigest()


def _grand(X, center, x, mu) -> np.ndarray:
    """6 x 6 grand resistance from the (3N, 6) forces of the six modes."""
    f = x.reshape(len(X), 3, 6).transpose(2, 0, 1) * float(np.real(mu))
    r = X - np.asarray(center, float)
    R = np.empty((6, 6))
    R[:3] = f.sum(axis=1).T
    R[3:] = np.cross(r[None], f).sum(axis=1).T
    return R


def solve_resistance(payload: dict, precision: str, cache: FactorCache | None, stream=None, graphs=None):
    """Six-mode grand resistance of one rigid body: factor (or reuse the
    resident factors) and solve the six rigid-body right-hand sides.  With a
    ``graphs`` pool (graphs.GraphPool), mixed-precision solves of at most
    graphs.GRAPH_MAX_POINTS points run as a captured graph on a padded
    system (resident factors then stay padded)."""
    from . import linalg
    from .core.solve import rigid_modes

    t = D.torch()
    kernel, X, center, eps = payload["kernel"], payload["points"], payload["center"], payload["eps"]
    dev = D.device()
    if graphs is not None and precision == "mixed":
        from . import graphs as G
        if len(X) <= G.GRAPH_MAX_POINTS and graphs.members > 1:
            R, err = solve_resistance_group([payload], cache, stream, graphs)[0]
            if err is not None:
                raise RuntimeError(err)
            return R
        if len(X) <= G.GRAPH_MAX_POINTS:
            key = geometry_key(kernel, X, eps) + "graph" if cache is not None else None
            item = cache.get(key) if cache is not None else None
            modes = rigid_modes(X, center)
            if item is None:
                g = graphs.get(len(X), stream)
                x, _, F = g.solve(kernel, X, eps, modes)
                if cache is not None:
                    K64p, Fp = g.resident(F)
                    cache.put(key, (K64p, Fp), K64p.numel() * 8 + Fp.nbytes())
            else:
                x, _ = G.solve_resident(item[0], item[1], modes, steps=graphs.steps, stream=stream)
            return _grand(X, center, x, payload.get("mu", 1.0))
    # larger solves: identity-padded to a multiple of 128 rows
    from .core.assemble import assemble_padded
    key = geometry_key(kernel, X, eps) + precision if cache is not None else None
    item = cache.get(key) if cache is not None else None
    modes = rigid_modes(X, center)
    n = 3 * len(X)
    if item is None:
        K64, K32, b, _ = assemble_padded(kernel, X, eps, modes, fp32=(precision == "mixed"), stream=stream)
        if precision == "mixed":
            # one blob width: a symmetric matrix, factored as L D L^T
            F = linalg.getrf(K32, stream=stream, refill=lambda A: A.copy_(K64), products="bf16x3",
                             sym=np.ndim(eps) == 0 or np.size(eps) == 1)
        else:
            F = linalg.getrf(K64.clone(), stream=stream, refill=lambda A: A.copy_(K64))
        if F.info > 0:
            raise np.linalg.LinAlgError("Matrix is singular.")
        item = (K64, F)
        if cache is not None:
            cache.put(key, item, K64.numel() * 8 + F.nbytes())
    else:
        with t.cuda.stream(stream) if stream is not None else nullcontext():
            b = t.zeros((int(item[0].shape[0]), 6), dtype=t.float64, device=dev)
            b[:n].copy_(t.from_numpy(modes))
    K64, F = item[0], item[1]
    if precision == "mixed":
        x, steps, _ = linalg.refine(K64, F, b, tol=1e-10, stream=stream)
    else:
        x = b
        linalg.getrs(F, x, stream=stream)
    return _grand(X, center, x[:n].cpu().numpy(), payload.get("mu", 1.0))


def solve_resistance_group(payloads: list, cache: FactorCache | None, stream, graphs) -> list:
    """Six-mode grand resistances of several rigid bodies of one graph
    bucket, as solve_resistance computes each alone (bit for bit): cached
    geometries are solved on their resident factors, the others together in
    one grouped graph launch.  Returns (R, None) or (None, error text) per
    payload."""
    from . import graphs as G
    from .core.solve import rigid_modes

    out: list = [None] * len(payloads)
    todo = []
    for i, p in enumerate(payloads):
        X = p["points"]
        modes = rigid_modes(X, p["center"])
        key = geometry_key(p["kernel"], X, p["eps"]) + "graph" if cache is not None else None
        item = cache.get(key) if cache is not None else None
        if item is not None:
            try:
                x, _ = G.solve_resident(item[0], item[1], modes, steps=graphs.steps, stream=stream)
                out[i] = (_grand(X, p["center"], x, p.get("mu", 1.0)), None)
            except Exception as exc:  # reported per request
                out[i] = (None, f"{type(exc).__name__}: {exc}")
        else:
            todo.append((i, key, modes))
    if todo:
        g = graphs.get_group(len(payloads[todo[0][0]]["points"]), stream)
        probs = [(payloads[i]["kernel"], payloads[i]["points"], payloads[i]["eps"], modes) for i, _, modes in todo]
        res = g.solve_many(probs, raise_errors=False)
        for z, ((i, key, _), r) in enumerate(zip(todo, res)):
            if isinstance(r, Exception):
                out[i] = (None, f"{type(r).__name__}: {r}")
                continue
            x, _, F = r
            if cache is not None:
                K64p, Fp = g.resident(z, F)
                cache.put(key, (K64p, Fp), K64p.numel() * 8 + Fp.nbytes())
            p = payloads[i]
            out[i] = (_grand(p["points"], p["center"], x, p.get("mu", 1.0)), None)
    return out


def new_session(spec: dict, precision: str):
    """The swimmer session a trace's swim_step requests advance (created at
    its first step)."""
    from .swimming import SwimmerSession
    return SwimmerSession(spec["shape_at"], spec["dt"], spec["eps"], spec.get("mu", 1.0), spec.get("x0", (0, 0, 0)),
                          spec.get("B0"), spec.get("kernel"), spec.get("t0", 0.0), precision=precision)


def run_swimmer(payload: dict, precision: str, cache: FactorCache | None = None, stream=None):
    from .swimming import swim_rk2
    tr = swim_rk2(payload["shape_at"], 0.0, payload["dt"], payload["steps"], payload["eps"],
                  precision=precision, stream=stream, factor_cache=cache)
    return np.concatenate([tr.x0[-1], tr.B[-1].ravel()])


class SolveServer:
    """Continuous-batching server: `slots` concurrent requests, each on its
    own stream, admitted by HBM footprint against `budget_bytes`.  With
    ``graphs`` (mixed precision) small resistance solves replay captured
    graphs, whose buffers (at most graphs.pool_bytes(slots)) are set aside
    from the budget.  The graphs are captured here, in the constructing
    thread: build the server while no other thread of the process is
    using the GPU (a capture that overlaps other threads' runtime calls
    can end the process).

    With ``group`` > 1 (graphs on), a slot that takes a small resistance
    request also takes up to group - 1 other waiting small resistance
    requests of the same graph bucket (the earliest arrived) and solves them
    together in one grouped graph launch (graphs.GroupSolveGraph); each
    result is bit for bit the one it gets alone."""

    def __init__(self, slots: int = 4, precision: str = "mixed", budget_bytes: int | None = None,
                 cache_bytes: int = 0, graphs: bool = True, group: int = 1):
        t = D.torch()
        dev = D.device()
        if budget_bytes is None:
            free, _ = t.cuda.mem_get_info(dev)
            budget_bytes = int(0.7 * free)
        self.slots = slots
        self.precision = precision
        self.budget = budget_bytes
        self.cache = FactorCache(cache_bytes) if cache_bytes > 0 else None
        self.budget -= max(0, int(cache_bytes))            # resident factors live inside the budget
        # one stream per slot, created for this server (never shared with
        # another user of the process), and one long-lived worker thread per
        # slot that takes its requests from a queue
        self.streams = [D.lease_stream() for _ in range(slots)]
        self.graphs = None
        self.group = 1
        if graphs and precision == "mixed":
            from .graphs import GraphPool, pool_bytes
            self.group = max(1, int(group))
            # every capture here, before any worker thread
            self.graphs = GraphPool(streams=self.streams, members=self.group)
            self.budget -= pool_bytes(slots, self.group)
        from .graphs import GRAPH_MAX_POINTS
        self.graphs_max_points = GRAPH_MAX_POINTS
        self._jobs = [queue.SimpleQueue() for _ in range(slots)]
        self._workers = [threading.Thread(target=self._worker, args=(sid,), daemon=True,
                                          name=f"fl-slot-{sid}") for sid in range(slots)]
        for w in self._workers:
            w.start()

    def _worker(self, sid: int):
        dev = D.device()
        t = D.torch()
        t.cuda.set_device(dev)
        while True:
            job = self._jobs[sid].get()
            if job is None:
                return
            job()

    def close(self) -> None:
        """Stop the slot workers (after the requests they hold), drop the
        resident factors and the captured graphs (their device memory goes
        back to the allocator; the cache's hit and miss counts stay) and hand
        the slots' streams back."""
        if not self._workers:
            return
        for q in self._jobs:
            q.put(None)
        for w in self._workers:
            w.join()
        self._workers = []
        for s in self.streams:
            s.synchronize()
            D.release_stream(s)
        self.streams = []
        if self.cache is not None:
            self.cache.clear()
        self.graphs = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _execute(self, req: Request, stream, session=None):
        t = D.torch()
        with t.cuda.stream(stream):
            if req.kind == "resistance":
                out = solve_resistance(req.payload, self.precision, self.cache, stream=stream, graphs=self.graphs)
            elif req.kind == "swimmer":
                out = run_swimmer(req.payload, self.precision, self.cache, stream=stream)
            elif req.kind == "swim_step":
                out = session.step(stream=stream, factor_cache=self.cache)
            else:
                raise ValueError(f"unknown request kind {req.kind!r}")
            stream.synchronize()
            return out

    def serve(self, requests: list[Request], realtime: bool = True, on_result=None) -> list[Result]:
        """Serve a trace.  With realtime=True requests become visible at their
        arrival times (Poisson traces); otherwise all are queued at t = 0.
        ``on_result(result)``, if given, is called from the worker thread as
        each request completes (not for requests rejected at admission)."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        results: list[Result] = []
        lock = threading.Condition()
        state = {"in_use": 0, "running": 0, "t0": None, "next": 0}
        free_streams = list(range(self.slots))
        waiting: list = []
        # swimmer sessions of this trace: the steps of one session run one at
        # a time, in arrival order; only a session's next step is a candidate
        sessions: dict = {}          # key -> SwimmerSession (made at its first step)
        pending: dict = {}           # key -> arrived steps not yet admitted, in order
        busy: set = set()            # sessions with a step queued for admission or running

        def now():
            return time.perf_counter() - state["t0"]

        def offer_next_step(key):
            # under lock: the session's next arrived step becomes a candidate
            if key not in busy and pending.get(key):
                nxt = pending[key].pop(0)
                busy.add(key)
                heapq.heappush(waiting, (nxt.arrival, nxt.rid, nxt))

        def worker_body(batch: list, sid: int, fp: int):
            start = now()
            req = batch[0]
            if len(batch) > 1:
                st = self.streams[sid]
                try:
                    t = D.torch()
                    with t.cuda.stream(st):
                        outs = solve_resistance_group([r.payload for r in batch], self.cache, st, self.graphs)
                except Exception as exc:  # the whole group failed: each request reports it
                    outs = [(None, f"{type(exc).__name__}: {exc}")] * len(batch)
            else:
                err, value = None, None
                try:
                    session = None
                    if req.kind == "swim_step":
                        key = req.payload["session"]
                        if key not in sessions:
                            sessions[key] = new_session(req.payload["spec"], self.precision)
                        session = sessions[key]
                    value = self._execute(req, self.streams[sid], session)
                except Exception as exc:  # a failing request is reported, the server keeps running
                    err = f"{type(exc).__name__}: {exc}"
                outs = [(value, err)]
            end = now()
            done = [Result(r.rid, r.kind, v, r.arrival if realtime else 0.0, start, end, e)
                    for r, (v, e) in zip(batch, outs)]
            with lock:
                results.extend(done)
                state["in_use"] -= fp
                state["running"] -= 1
                free_streams.append(sid)
                if req.kind == "swim_step":
                    busy.discard(req.payload["session"])
                    offer_next_step(req.payload["session"])
                lock.notify_all()
            if on_result is not None:
                for res in done:
                    on_result(res)

        def groupable(r: Request) -> bool:
            return (self.group > 1 and r.kind == "resistance"
                    and len(r.payload["points"]) <= self.graphs_max_points)

        def take_group(head: Request, fp: int):
            # under lock: the head request and up to group - 1 waiting requests
            # of its bucket, earliest first, while their footprints fit
            from .graphs import bucket_rows
            if not groupable(head):
                return [head], fp
            rows = bucket_rows(len(head.payload["points"]))
            same = sorted(e for e in waiting if groupable(e[2]) and bucket_rows(len(e[2].payload["points"])) == rows)
            batch = [head]
            for e in same[:self.group - 1]:
                f = footprint_bytes(e[2], self.precision, shared_cache=self.cache is not None)
                if state["in_use"] + fp + f > self.budget:
                    break
                batch.append(e[2])
                fp += f
            if len(batch) > 1:
                taken = {r.rid for r in batch[1:]}
                waiting[:] = [e for e in waiting if e[2].rid not in taken]
                heapq.heapify(waiting)
            return batch, fp

        if not self._workers:
            raise RuntimeError("the server is closed")
        state["t0"] = time.perf_counter()
        i = 0
        with lock:
            while i < len(reqs) or waiting or state["running"]:
                # move arrived requests into the waiting list (arrival order)
                while i < len(reqs) and (not realtime or reqs[i].arrival <= now()):
                    r = reqs[i]
                    if r.kind == "swim_step":
                        pending.setdefault(r.payload["session"], []).append(r)
                        offer_next_step(r.payload["session"])
                    else:
                        heapq.heappush(waiting, (r.arrival, r.rid, r))
                    i += 1
                admitted = False
                while waiting and free_streams:
                    req = waiting[0][2]
                    fp = footprint_bytes(req, self.precision, shared_cache=self.cache is not None)
                    if fp > self.budget:
                        heapq.heappop(waiting)
                        results.append(Result(req.rid, req.kind, None, req.arrival if realtime else 0.0, now(), now(),
                                              f"rejected: footprint {fp} B exceeds the budget {self.budget} B"))
                        if req.kind == "swim_step":
                            busy.discard(req.payload["session"])
                            offer_next_step(req.payload["session"])
                        continue
                    if state["in_use"] + fp > self.budget and state["running"] > 0:
                        break  # wait for memory to free up
                    heapq.heappop(waiting)
                    batch, fp = take_group(req, fp)
                    sid = free_streams.pop()
                    state["in_use"] += fp
                    state["running"] += 1
                    self._jobs[sid].put(lambda batch=batch, sid=sid, fp=fp: worker_body(batch, sid, fp))
                    admitted = True
                if admitted:
                    continue
                timeout = None
                if realtime and i < len(reqs):
                    timeout = max(0.0, reqs[i].arrival - now())
                lock.wait(timeout=timeout if timeout is not None else 1.0)
        return sorted(results, key=lambda r: r.rid)


def summarize(results: list[Result]) -> dict:
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
    span = max((r.end for r in results), default=0.0)
    return {
        "requests": len(results),
        "completed": len(ok),
        "failed": len(results) - len(ok),
        "makespan_s": span,
        "solves_per_s": len(ok) / span if span > 0 else 0.0,
        "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "latency_mean_s": float(lat.mean()) if len(lat) else None,
        # the three slowest requests: (request id, kind, wait before start, latency)
        "slowest": [(r.rid, r.kind, round(r.start - r.arrival, 4), round(r.latency, 4))
                    for r in sorted(ok, key=lambda r: -r.latency)[:3]],
    }


def serving_mix(n_requests: int, seed: int = 5, n_min: int = 200, n_max: int = 6000, swimmer_frac: float = 0.3,
                swimmer_steps: int = 200, swimmer_body: int = 1000, swimmer_flag: int = 400,
                rate: float = 1.0, swimmer_mode: str = "steps") -> list[Request]:
    """The serving workload: Poisson arrivals at `rate` per second; 70 %
    rigid-body resistance solves on seeded ellipsoids (semi-axes U[0.5, 2],
    N uniform in [n_min, n_max], Fibonacci points mapped onto the ellipsoid,
    centre height U[1.2, 4], draws whose vertical semi-axis reaches within
    eps of the wall rejected; WallStokeslet3D or free-space Brinkmanlet3D with
    alpha ~ U[0, 10], 50/50; eps = 0.5 sqrt(area / N); six rigid-body
    right-hand sides) and 30 % biflagellate swimmer time steps (midpoint
    RK2, dt = one fiftieth of the beat period).

    swimmer_mode "steps": each swimmer request is one RK2 time step of a
    swimmer session whose state the server keeps between requests; the
    swimmer requests are dealt round-robin to round(n_swimmer / swimmer_steps)
    sessions, so each session runs about swimmer_steps steps (the 2,000
    requests of seed 5: 610 steps of 3 swimmers, 203-204 steps each).  The sessions start from seeded
    positions and orientations.  swimmer_mode "trajectory": each swimmer
    request runs a whole trajectory of swimmer_steps steps."""
    from .geometry import ellipsoid, ellipsoid_area
    from .kernels import Brinkmanlet3D, WallStokeslet3D
    from .swimming import biflagellate

    if swimmer_mode not in ("steps", "trajectory"):
        raise ValueError(f"unknown swimmer_mode {swimmer_mode!r}")
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    shape_at = None
    swim_rids = []
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if rng.random() < swimmer_frac:
            if shape_at is None:
                shape_at = biflagellate(n_body=swimmer_body, n_flag=swimmer_flag)
            swim_rids.append(rid)
            out.append(Request("swimmer", {"shape_at": shape_at, "dt": 1.0 / 50, "steps": swimmer_steps,
                                           "eps": 0.01, "n_points": swimmer_body + 2 * swimmer_flag}, t, rid))
            continue
        while True:
            axes = rng.uniform(0.5, 2.0, 3)
            N = int(rng.integers(n_min, n_max + 1))
            h = rng.uniform(1.2, 4.0)
            eps = 0.5 * np.sqrt(ellipsoid_area(axes) / N)
            if axes[2] < h - eps:
                break
        wall = rng.random() < 0.5
        if wall:
            kernel, center = WallStokeslet3D(), (0.0, 0.0, h)
        else:
            kernel, center = Brinkmanlet3D(rng.uniform(0.0, 10.0)), (0.0, 0.0, 0.0)
        X = ellipsoid(N, axes, center)
        out.append(Request("resistance", {"kernel": kernel, "points": X, "center": center, "eps": eps}, t, rid))
    if swimmer_mode == "steps" and swim_rids:
        n_sessions = max(1, int(round(len(swim_rids) / max(1, swimmer_steps))))
        srng = np.random.default_rng(seed + 1000)
        specs = []
        for k in range(n_sessions):
            q, _ = np.linalg.qr(srng.standard_normal((3, 3)))
            B0 = q * np.sign(np.linalg.det(q))
            specs.append({"shape_at": shape_at, "dt": 1.0 / 50, "eps": 0.01, "x0": srng.uniform(-5, 5, 3),
                          "B0": B0, "n_points": swimmer_body + 2 * swimmer_flag})
        for j, rid in enumerate(swim_rids):
            old = out[rid]
            k = j % n_sessions
            out[rid] = Request("swim_step", {"session": f"swimmer{k}", "spec": specs[k], "seq": j // n_sessions},
                               old.arrival, rid)
    return out
