#!/usr/bin/env python3
"""Benchmark of the MI355X kernel-sum hot path (one GPU).

Workload (the N-body kernel-sum drop-in): 262,144 points, sources = targets,
seeded uniform in [-4,4]^2 x [0.05,4] (numpy default_rng(2)), forces i.i.d.
N(0,1), eps = 0.02, mu = 1.  One step = one fl.evaluate of the fused
WallStokeslet3D kernel sum in FP64 over all 6.87e10 target-source pairs, with
inputs resident in HBM.  value = pair evaluations per second (M N / t).
Because the targets are the sources and the blob width is uniform, the
library takes its reciprocal path: the block of each pair in different
512-point tiles is evaluated once and applied to both points (G(y, x) =
G(x, y)^T); the roofline counts the flops that path performs.

Also reported in the same JSON line: the other kernels of the workload
(Stokeslet3D, Brinkmanlet3D(4), Stokeslet3D with mu = 1+0.5i) in FP64 and
FP32C with their errors against FP64, the roofline of the dominant kernel
(FP64 flops per pair / its average launch time, measured with HIP events on
the launch stream) and a CPU baseline: the oracle's C/OpenMP restatement of
the reference loop timed on this host's cores over a bounded sample.

The default run also carries the other legs of the workloads: the icosphere
solve, the 8,192-point sphere resistance solves (WallStokeslet3D and
Brinkmanlet3D at alpha = 1, 5, 10; mixed precision against this library's
FP64 LU and rocSOLVER's), the 131,072-point rod-carpet GMRES (FP64, flexible
with FP32C inner cycles, block-Jacobi), the 2,000-request serving mix and the
C2 kernel sum at 2,097,152 points by the treecode against the direct sum.
`bench_wall_s` is the whole run's wall time.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-solve] [--no-serve] [--no-gmres]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
N_POINTS = 262144
EPS = 0.02
SEED = 2
# FP64 flops per pair (add/sub/mul = 1, FMA = 2, sqrt/rsqrt/div/exp = 1;
# per-source constants amortised; real strengths).
FLOPS_PER_PAIR = {"wall": 116, "stokeslet": 32, "brinkman": 50}
# The transposed apply of the reciprocal (targets = sources) evaluation, by the
# same rules: w = d0 g0 + d1 g1 (3), v = b w + t g2 (3), three outputs (4 each).
FLOPS_TRANSPOSED_APPLY = {"wall": 18, "stokeslet": 17, "brinkman": 17}
SYM_TILE = 512


def progress(msg: str) -> None:
    """One line per finished section or pass, on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def sym_flops(n: int, fam: str) -> float:
    """Algorithmic flops of one reciprocal evaluation over n points: pairs in
    different 512-point tiles are evaluated once and applied both ways; pairs
    inside a tile (self-pairs included) are evaluated per ordered pair."""
    sizes = [min(SYM_TILE, n - k) for k in range(0, n, SYM_TILE)]
    inside = float(sum(m * m for m in sizes))
    across = (float(n) * n - inside) / 2
    return across * (FLOPS_PER_PAIR[fam] + FLOPS_TRANSPOSED_APPLY[fam]) + inside * FLOPS_PER_PAIR[fam]
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and FP64 matrix) peak
HBM_PEAK_GBS = 8000.0


def c2_inputs():
    rng = np.random.default_rng(SEED)
    X = rng.uniform([-4, -4, 0.05], [4, 4, 4], (N_POINTS, 3))
    F = rng.standard_normal((N_POINTS, 3))
    return X, F


def profile_traffic(name: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated),
    or None when no summary is present."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(X, F, n_targets: int):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O  # the oracle is the CPU baseline, not the product
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    O.set_threads(threads)
    O.evaluate("wall", X[:8], X[:2048], F[:2048], EPS)  # warm the library
    t0 = time.perf_counter()
    O.evaluate("wall", X[:n_targets], X, F, EPS)
    dt = time.perf_counter() - t0
    # the same loop on one core, on a sixteenth of the targets
    n1 = max(16, n_targets // 16)
    O.set_threads(1)
    t1 = time.perf_counter()
    O.evaluate("wall", X[:n1], X, F, EPS)
    dt1 = time.perf_counter() - t1
    O.set_threads(threads)
    return {"value": n_targets * N_POINTS / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
            "host_cores": cores, "value_1core": n1 * N_POINTS / dt1,
            "sample": f"WallStokeslet3D FP64, first {n_targets} targets x all {N_POINTS} sources "
                      f"({n_targets * N_POINTS:.3e} pairs, {dt:.1f} s; one core: {n1} targets, {dt1:.1f} s); "
                      f"C/OpenMP restatement of core/matvec.py:_apply with the reference's pow()-based block, "
                      f"-O3, no fast-math"}


def c1_section(torch, fl):
    """The reference's own CPU case: a level-3 icosphere (642 points),
    eps = 0.25 x mean nearest-neighbour spacing, Stokeslet3D, U = (1, 0, 0),
    dense LU (1926 x 1926), drag / (6 pi), then the forces evaluated at
    10,000 targets in the shell 1.5 < r < 5 (seed 1).  The same solve through
    scipy.linalg.solve (LAPACK, the reference's call at core/solve.py:66) is
    timed on the host with the kernel matrix this library assembled."""
    import scipy.linalg
    from fluidlets_amd.geometry import icosphere, mean_nearest_spacing

    X = icosphere(3)
    eps = 0.25 * mean_nearest_spacing(X)
    U = np.tile([1.0, 0.0, 0.0], (len(X), 1))
    rng = np.random.default_rng(1)
    d = rng.normal(size=(10000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    T = d * (rng.uniform(1.5 ** 3, 5.0 ** 3, 10000) ** (1.0 / 3.0))[:, None]
    K = fl.Stokeslet3D()
    fl.resistance(K, X, U, eps, method="lu")                       # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, info = fl.resistance(K, X, U, eps, method="lu")
    u = fl.evaluate(K, T, X, f, eps)
    t_gpu = time.perf_counter() - t0
    A = fl.assemble(K, X, X, eps)
    t1 = time.perf_counter()
    fc = scipy.linalg.solve(A, U.reshape(-1))
    t_cpu = time.perf_counter() - t1
    return {"points": len(X), "eps": eps, "drag_over_6pi": float(f[:, 0].sum() / (6 * np.pi)),
            "residual": info.residual, "gpu_solve_and_evaluate_s": t_gpu, "shell_velocity_max": float(np.abs(u).max()),
            "cpu_lapack_solve_s": t_cpu, "cpu_threads": len(os.sched_getaffinity(0)),
            "cpu_vs_gpu_force_relerr": float(np.linalg.norm(fc - f.reshape(-1)) / np.linalg.norm(fc))}


def c3_solves(torch, fl, D, kernels=("wall",), n_points=8192, repeats=2, vendor=True, cpu_lapack=True):
    """Mixed-precision wall resistance solve: Fibonacci sphere (radius 1,
    centre (0,0,1.5)), eps = 0.5 sqrt(4 pi / N), six rigid-body right-hand
    sides.  Times FP32-MFMA LU + FP64 refinement to 1e-10, the FP64 LU of this
    library and the FP64 LU of rocSOLVER (torch.linalg.lu_factor), each the
    best of `repeats` runs (the first run of each also pays library start-up),
    and reports each 6x6 grand-resistance entry's error against the FP64 LU."""
    from fluidlets_amd import linalg
    from fluidlets_amd.core.assemble import assemble_device
    from fluidlets_amd.core.solve import rigid_modes
    from fluidlets_amd.geometry import fibonacci_sphere

    def grand(X, c, x):
        f = x.reshape(len(X), 3, 6).transpose(2, 0, 1)
        r = X - np.asarray(c)
        R = np.zeros((6, 6))
        for m in range(6):
            R[:3, m] = f[m].sum(axis=0)
            R[3:, m] = np.cross(r, f[m]).sum(axis=0)
        return R

    out = {}
    N = n_points
    n = 3 * N
    h = np.sqrt(4 * np.pi / N)
    for name in kernels:
        if name == "wall":
            K, center = fl.WallStokeslet3D(), (0.0, 0.0, 1.5)
        else:
            K, center = fl.Brinkmanlet3D(float(name.split(":")[1])), (0.0, 0.0, 0.0)
        X = fibonacci_sphere(N, 1.0, center)
        dev = D.device()
        x_soa = D.soa(X, dev)
        eps_t, stride = D.eps_arg(0.5 * h, N, dev)
        rhs = torch.from_numpy(rigid_modes(X, center)).to(dev)
        ev = lambda: torch.cuda.Event(enable_timing=True)
        # assembly: FP64 + FP32 copy in one pass (HBM-write bound)
        assemble_device(K, x_soa, x_soa, eps_t, stride, 1.0, fp64=True, fp32=True)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        K64, K32 = assemble_device(K, x_soa, x_soa, eps_t, stride, 1.0, fp64=True, fp32=True)
        e1.record()
        torch.cuda.synchronize()
        asm_ms = e0.elapsed_time(e1)
        row = {"n": n, "assembly_ms": asm_ms, "assembly_GBps": 108.0 * N * N / (asm_ms * 1e-3) / 1e9}
        # plain FP64 LU (this library): factor + 6 solves
        best64 = None
        for _ in range(repeats):
            A = K64.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            F = linalg.getrf(A, refill=lambda X: X.copy_(K64))
            x64 = rhs.clone()
            linalg.getrs(F, x64)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best64 = dt if best64 is None else min(best64, dt)
            del A
        row["fp64_lu_s"] = best64
        row["fp64_general_panels"] = F.general_panels
        row["fp64_lu_async"] = bool(F.extra.get("async", False))
        x64h = x64.cpu().numpy()
        del F
        # mixed: FP32 LU + FP64 refinement (the FP32 copy is refactored each
        # repeat).  The library's mixed solves take split-bf16 trailing
        # products ("bf16x3"); FP32 MFMA products are timed beside them.
        # residual tolerance 1e-12: the 1e-10 residual the workload asks for
        # leaves the forces' error against FP64 near 1e-10 on this matrix
        # (condition ~70); one more correction puts it near 1e-14
        refine_tol = 1e-12
        row["refinement_tol"] = refine_tol
        # the symmetric factorization (A = L D L^T, half the flops; these
        # matrices are symmetric by reciprocity), FP32 and split-bf16
        # products, then the LU ones; the LU with split-bf16 products last
        # (its numbers are the row's headline mixed solve)
        for sym, products in ((True, "fp32"), (True, "bf16x3"), (False, "fp32"), (False, "bf16x3")):
            bestm = None
            for _ in range(repeats):
                A32 = K32.clone()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                F32 = linalg.getrf(A32, refill=lambda X: X.copy_(K32), products=products, sym=sym)
                xm, steps, res = linalg.refine(K64, F32, rhs, tol=refine_tol)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                bestm = dt if bestm is None else min(bestm, dt)
                t1 = time.perf_counter()
                xm2, _, _ = linalg.refine(K64, F32, rhs, tol=refine_tol)
                torch.cuda.synchronize()
                resolve = time.perf_counter() - t1
                sym_ran = bool(F32.extra.get("sym"))
                del F32, A32
            xmh = xm.cpu().numpy()
            err = float(np.linalg.norm(xmh - x64h) / np.linalg.norm(x64h))
            tag = ("mixed_sym_" if sym else "mixed_") + products + "_products"
            if sym or products == "fp32":
                row[tag + "_lu_refine_s"] = bestm
                row[tag + "_refinement_steps"] = steps
                row[tag + "_vs_fp64_force_relerr"] = err
                row[tag + "_speedup_vs_fp64_lu"] = best64 / bestm
            if sym:
                row[tag + "_factorization_was_symmetric"] = sym_ran
        row["mixed_products"] = "bf16x3"
        row["mixed_lu_refine_s"] = bestm
        row["mixed_resolve_s"] = resolve
        row["refinement_steps"] = steps
        row["refinement_residual_max"] = float(res.max())
        row["mixed_vs_fp64_force_relerr"] = err
        R64, Rm = grand(X, center, x64h), grand(X, center, xmh)
        row["grand_resistance_fp64"] = R64.round(6).tolist()
        row["grand_resistance_max_entry_relerr"] = float(np.abs(Rm - R64).max() / np.abs(R64).max())
        row["speedup_vs_fp64_lu"] = best64 / bestm
        row["best_mixed_s"] = min(bestm, row.get("mixed_sym_bf16x3_products_lu_refine_s", bestm))
        row["best_speedup_vs_fp64_lu"] = best64 / row["best_mixed_s"]
        # rocSOLVER FP64 (torch.linalg) as the library baseline, best of `repeats` like the others
        try:
            if not vendor:
                raise RuntimeError("vendor LU skipped (--no-vendor-lu)")
            bestr = None
            for _ in range(repeats):
                A = K64.clone()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                LUt, piv = torch.linalg.lu_factor(A)
                xr = torch.linalg.lu_solve(LUt, piv, rhs)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                bestr = dt if bestr is None else min(bestr, dt)
                if _ + 1 < repeats:
                    del A, LUt, piv, xr
            row["rocsolver_fp64_lu_s"] = bestr
            row["speedup_vs_rocsolver_fp64"] = row["rocsolver_fp64_lu_s"] / bestm
            row["rocsolver_vs_fp64_force_relerr"] = float(torch.linalg.norm(xr - x64) / torch.linalg.norm(x64))
            del A, LUt, xr
        except Exception as exc:  # report, do not fail the bench
            row["rocsolver_error"] = repr(exc)[:200]
        flops = 2.0 * n**3 / 3
        row["fp64_lu_tflops"] = flops / best64 / 1e12
        row["mixed_lu_tflops_equiv"] = flops / bestm / 1e12
        row["solves_per_s_mixed"] = 1.0 / bestm
        if cpu_lapack and name == kernels[0]:
            # the reference's dense solve, scipy.linalg.solve (LAPACK getrf + getrs, threads = host cores)
            import scipy.linalg
            try:
                Kh = K64.cpu().numpy()
                bh = rhs.cpu().numpy()
                t1 = time.perf_counter()
                xc = scipy.linalg.solve(Kh, bh, overwrite_a=True, check_finite=False)
                row["cpu_lapack_fp64_solve_s"] = time.perf_counter() - t1
                row["cpu_lapack_threads"] = len(os.sched_getaffinity(0))
                row["cpu_lapack_vs_fp64_force_relerr"] = float(np.linalg.norm(xc - x64h) / np.linalg.norm(x64h))
                row["speedup_mixed_vs_cpu_lapack"] = row["cpu_lapack_fp64_solve_s"] / bestm
                del Kh, xc
            except Exception as exc:  # report, do not fail the bench
                row["cpu_lapack_error"] = repr(exc)[:200]
        out[name] = row
        progress(f"wall-sphere solve {name}: mixed LU {row['mixed_lu_refine_s']:.4f} s, mixed symmetric "
                 f"{row['mixed_sym_bf16x3_products_lu_refine_s']:.4f} s (FP32 products: LU "
                 f"{row['mixed_fp32_products_lu_refine_s']:.4f} s, symmetric "
                 f"{row['mixed_sym_fp32_products_lu_refine_s']:.4f} s), FP64 LU {row['fp64_lu_s']:.4f} s")
        del K64, K32
        torch.cuda.empty_cache()
    return out


def serving_section(torch, n_requests=2000, baseline_requests=300, swimmer_steps=200, n_max=6000, slots=4, seed=5,
                    cache_gib=48, load=0.9, small_solves=True, group=8):
    """Serving mix: n_requests Poisson-arrival requests, 70 % six-mode
    resistance solves on ellipsoids (N uniform in [200, n_max]) and 30 %
    biflagellate swimmer time steps (RK2 steps of swimmer sessions of about
    swimmer_steps steps each, their state kept by the server between steps).

    - The continuous-batching server (mixed precision, `slots` streams,
      small solves grouped up to `group` per launch, a resident factor
      cache of cache_gib GiB that starts empty) serves the
      whole trace offline (all queued at t = 0: its capacity) and online
      (Poisson arrivals at `load` x that capacity).
    - The baseline, one request at a time in plain FP64 on the same GPU
      (every solve factors its matrix), serves the first baseline_requests
      requests of the trace (the same mix) offline and online at `load` x
      its own capacity; the server then serves that same subset at the
      baseline's online arrival rate, for latencies at equal load."""
    from fluidlets_amd.serving import Request, SolveServer, serving_mix, summarize

    reqs = serving_mix(n_requests, seed=seed, n_max=n_max, swimmer_steps=swimmer_steps, rate=1.0)
    sub = sorted(reqs, key=lambda r: r.arrival)[:baseline_requests]
    warm = [r for r in reqs if r.kind == "resistance"][:2]
    base_srv = SolveServer(slots=1, precision="fp64")
    base_srv.serve(warm, realtime=False)               # warm-up: first allocations and library start-up
    with SolveServer(slots=slots, precision="mixed", group=group) as w:
        w.serve(warm, realtime=False)

    def fresh():
        return SolveServer(slots=slots, precision="mixed", cache_bytes=int(cache_gib) << 30, group=group)

    def tick(tag, total):
        done = [0]

        def on_result(res):
            done[0] += 1
            if done[0] % 100 == 0 or done[0] == total:
                progress(f"serving: {tag} {done[0]}/{total} done")
        return on_result

    def poisson(trace, lam, rs):
        rng = np.random.default_rng(rs)
        t, out = 0.0, []
        for r in sorted(trace, key=lambda r: r.arrival):
            t += rng.exponential(1.0 / lam)
            out.append(Request(r.kind, r.payload, t, r.rid))
        return out

    def served(srv, trace, realtime, tag):
        res = srv.serve(trace, realtime=realtime, on_result=tick(tag, len(trace)))
        s = summarize(res)
        errors = [r.error for r in res if r.error]
        if errors:
            s["first_errors"] = errors[:3]
            progress(f"serving: {tag}: {len(errors)} of {len(res)} requests failed, first: {errors[0]}")
        if srv.cache is not None:
            s["cache_hits"], s["cache_misses"] = srv.cache.hits, srv.cache.misses
        torch.cuda.empty_cache()
        return s

    def fmt(v, f=".4f"):
        return "n/a" if v is None else format(v, f)

    kinds = {k: sum(r.kind == k for r in reqs) for k in ("resistance", "swim_step")}
    sessions = sorted({r.payload["session"] for r in reqs if r.kind == "swim_step"})
    out = {"requests": n_requests, "kinds": kinds, "swimmer_sessions": len(sessions),
           "steps_per_session": [sum(r.kind == "swim_step" and r.payload["session"] == k for r in reqs)
                                 for k in sessions],
           "slots": slots, "n_max": n_max, "server_cache_gib": cache_gib, "load": load,
           "baseline_requests": len(sub)}
    t_all = time.perf_counter()
    with fresh() as srv:
        off_fast = served(srv, reqs, False, "offline server")
    progress(f"serving: offline server {off_fast['solves_per_s']:.2f} requests/s")
    off_base = served(base_srv, sub, False, "offline baseline")
    progress(f"serving: offline baseline {off_base['solves_per_s']:.3f} requests/s")
    lam_s = load * off_fast["solves_per_s"]
    with fresh() as srv:
        on_fast = served(srv, poisson(reqs, lam_s, seed + 1), True, "online server")
    progress(f"serving: online server at {lam_s:.1f}/s: p50 {fmt(on_fast['latency_p50_s'])} s, "
             f"p99 {fmt(on_fast['latency_p99_s'])} s")
    lam_b = load * off_base["solves_per_s"]
    sub_trace = poisson(sub, lam_b, seed + 2)
    on_base = served(base_srv, sub_trace, True, "online baseline")
    progress(f"serving: online baseline at {lam_b:.2f}/s: p99 {fmt(on_base['latency_p99_s'])} s")
    with fresh() as srv:
        on_fast_sub = served(srv, sub_trace, True, "server at the baseline's rate")
    base_srv.close()
    out["offline_server_mixed"] = off_fast
    out["offline_baseline_fp64_one_at_a_time"] = off_base
    out["offline_throughput_speedup"] = off_fast["solves_per_s"] / off_base["solves_per_s"]
    out["online_server_mixed"] = dict(on_fast, arrival_rate_per_s=lam_s)
    out["online_baseline_fp64_one_at_a_time"] = dict(on_base, arrival_rate_per_s=lam_b)
    out["online_server_mixed_at_baseline_rate"] = dict(on_fast_sub, arrival_rate_per_s=lam_b)
    out["wall_s"] = time.perf_counter() - t_all
    if small_solves:
        out["small_solves"] = small_solves_section(slots=slots, seed=seed)
    return out


def small_solves_section(slots=4, seed=5, n_requests=192, passes=3, modes=("grouped", "graphs", "direct"),
                         group=8):
    """Small six-mode resistance solves (N uniform in [200, 640] points, the
    serving mix's ellipsoids and kernels), all queued at t = 0, served with
    grouped graphs (up to `group` requests of one bucket per launch), with
    one captured graph per request (padded buckets, fixed corrections) and
    without graphs (unpadded, adaptive refinement), on one slot and on
    `slots` slots.  Each server serves the trace once to warm up
    (allocations), then `passes` times timed; the pass with the median
    throughput is reported.  ``modes`` selects the paths."""
    from fluidlets_amd.serving import SolveServer, serving_mix, summarize

    reqs = serving_mix(n_requests, seed=seed + 2, n_min=200, n_max=640, swimmer_frac=0.0, rate=1.0)
    out = {"requests": n_requests, "n_points": [200, 640], "passes": passes}
    out["group"] = group
    for k in sorted({1, slots}):
        for mode in modes:
            t0 = time.perf_counter()
            srv = SolveServer(slots=k, precision="mixed", graphs=mode != "direct",
                              group=group if mode == "grouped" else 1)
            build_s = time.perf_counter() - t0
            srv.serve(reqs, realtime=False)
            runs = sorted((summarize(srv.serve(reqs, realtime=False)) for _ in range(passes)),
                          key=lambda s: s["solves_per_s"])
            s = runs[len(runs) // 2]
            tag = f"{mode}_{k}_slots"
            out[tag] = {key: s[key] for key in ("solves_per_s", "latency_p50_s", "latency_p99_s", "makespan_s")}
            out[tag]["solves_per_s_range"] = [runs[0]["solves_per_s"], runs[-1]["solves_per_s"]]
            if mode != "direct":
                out[tag]["graph_captures"] = srv.graphs.captures
                out[tag]["server_build_s"] = build_s
            srv.close()
            progress(f"serving: small solves {tag} {s['solves_per_s']:.1f} requests/s")
    for k in sorted({1, slots}):
        for mode in ("graphs", "grouped"):
            if f"{mode}_{k}_slots" in out and f"direct_{k}_slots" in out:
                out[f"{mode}_speedup_{k}_slots"] = (out[f"{mode}_{k}_slots"]["solves_per_s"]
                                                    / out[f"direct_{k}_slots"]["solves_per_s"])
    return out


def gmres_section(torch, fl, D, precision_list=("fp64", "mixed", "fp64_block_jacobi"), n_side=32, per_rod=128,
                  restart=50, rtol=1e-8, maxiter=40):
    """Matrix-free GMRES(restart) on the fixed-rod carpet: 32 x 32 rods of
    128 helical-lattice points (131,072 points), WallStokeslet3D, eps = 0.01,
    rods held fixed in the background shear u = (z, 0, 0) (right-hand side
    -(z, 0, 0)); FP64 GMRES, flexible GMRES with FP32C inner cycles, and FP64
    GMRES preconditioned by the rods' diagonal blocks (block-Jacobi, setup
    included in the time)."""
    from fluidlets_amd.core.gmres import gmres_device
    from fluidlets_amd.geometry import rod_carpet

    X = rod_carpet(n_side=n_side, per_rod=per_rod)
    N = len(X)
    dev = D.device()
    x_soa = D.soa(X, dev)
    eps_t, stride = D.eps_arg(0.01, N, dev)
    U = np.zeros((N, 3))
    U[:, 0] = -X[:, 2]
    U_soa = D.soa(U, dev)
    out = {"points": N, "restart": restart, "rtol": rtol}
    ref = None
    for prec in precision_list:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bj = prec == "fp64_block_jacobi"
        f, its, res = gmres_device(fl.WallStokeslet3D(), x_soa, eps_t, stride, U_soa, rtol=rtol, restart=restart,
                                   maxiter=maxiter, precision="fp64" if bj else prec,
                                   block_points=per_rod if bj else 0)
        torch.cuda.synchronize()
        row = {"seconds": time.perf_counter() - t0, "iterations": its, "residual": res}
        if ref is None:
            ref = f.clone()
        else:
            row["force_relerr_vs_fp64"] = (torch.linalg.norm(f - ref) / torch.linalg.norm(ref)).item()
        out[prec] = row
    return out


def treecode_section(torch, fl, D, n_points=2097152, degree=8, theta=0.7, leaf=1024, check_targets=8192):
    """The C2 kernel sum (WallStokeslet3D, eps = 0.02, points uniform in
    [-4,4]^2 x [0.05,4], N(0,1) forces; seed 3) at 2,097,152 points, by the
    barycentric Lagrange treecode (degree 8, theta 0.7, leaves of 2,048) and by
    one direct FP64 launch (the general path: above about 428,000 points the
    reciprocal path's workspace passes its 4 GiB cap).  The treecode's error is
    measured against the direct sum on 8,192 of the targets; its setup (trees
    and lists on the host, uploads) is reported apart from the evaluation."""
    from fluidlets_amd.core.matvec import evaluate_device
    from fluidlets_amd.treecode import Treecode

    rng = np.random.default_rng(3)
    X = rng.uniform([-4, -4, 0.05], [4, 4, 4], (n_points, 3))
    F = rng.standard_normal((n_points, 3))
    K = fl.WallStokeslet3D()
    dev = D.device()
    x, f = D.soa(X, dev), D.soa(F, dev)
    t0 = time.perf_counter()
    tc = Treecode(K, X, X, EPS, degree=degree, theta=theta, leaf_size=leaf)
    setup = time.perf_counter() - t0
    tc.apply_device(f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, _ = tc.apply_device(f)
    torch.cuda.synchronize()
    t_tc = time.perf_counter() - t0
    eps_t, stride = D.eps_arg(EPS, n_points, dev)
    t0 = time.perf_counter()
    ud_full, _ = evaluate_device(K, x, x, eps_t, stride, f)
    torch.cuda.synchronize()
    t_dir = time.perf_counter() - t0
    sub = np.arange(0, n_points, n_points // check_targets)[:check_targets]
    ud = ud_full[:, torch.from_numpy(sub).to(dev)].cpu().numpy()
    ut = u[:, torch.from_numpy(sub).to(dev)].cpu().numpy()
    del ud_full
    return {"points": n_points, "kernel": "WallStokeslet3D", "degree": degree, "theta": theta, "leaf_size": leaf,
            "setup_s": setup, "treecode_s": t_tc, "direct_s": t_dir, "speedup_vs_direct": t_dir / t_tc,
            "rel_err_vs_direct": float(np.linalg.norm(ut - ud) / np.linalg.norm(ud)),
            "checked_targets": int(len(sub)), "clusters_with_proxies": tc.ncl,
            "direct_pairs_frac": tc.n_direct_pairs / n_points**2, "proxy_pairs_frac": tc.n_proxy_pairs / n_points**2,
            "pair_evals_per_s": (tc.n_direct_pairs + tc.n_proxy_pairs) / t_tc}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-targets", type=int, default=16384)
    ap.add_argument("--no-table", action="store_true", help="skip the other kernels of the workload")
    ap.add_argument("--no-solve", action="store_true", help="skip the N=8192 resistance-solve section")
    ap.add_argument("--solve-kernels", default="wall,brinkman:1,brinkman:5,brinkman:10",
                    help="comma list: wall, brinkman:1, brinkman:5, ...")
    ap.add_argument("--solve-points", type=int, default=8192)
    ap.add_argument("--no-vendor-lu", action="store_true", help="skip the rocSOLVER FP64 LU comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-mix section")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-baseline-requests", type=int, default=300,
                    help="requests of the trace (its first, in arrival order) the one-at-a-time baseline serves")
    ap.add_argument("--serve-swimmer-steps", type=int, default=200)
    ap.add_argument("--serve-cache-gib", type=int, default=48, help="resident factor cache of the mixed server")
    ap.add_argument("--serve-nmax", type=int, default=6000)
    ap.add_argument("--serve-load", type=float, default=0.9, help="online arrival rate as a fraction of capacity")
    ap.add_argument("--small-solves-only", action="store_true",
                    help="of the serving section, run only the small-solve comparison")
    ap.add_argument("--no-gmres", action="store_true", help="skip the 131,072-point GMRES rod-carpet section")
    ap.add_argument("--no-treecode", action="store_true", help="skip the 2,097,152-point treecode section")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (the reference point for HBM byte counters)")
    args = ap.parse_args()
    t_bench = time.perf_counter()
    if args.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(the kernel sum shards by targets with an all-gather of sources; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import fluidlets_amd as fl
    from fluidlets_amd import _device as D
    from fluidlets_amd.core.matvec import evaluate_device

    dev = D.device()
    X, F = c2_inputs()
    x = D.soa(X, dev)
    f = D.soa(F, dev)
    eps_t, stride = D.eps_arg(EPS, N_POINTS, dev)
    wall = fl.WallStokeslet3D()
    u = torch.empty((3, N_POINTS), dtype=torch.float64, device=dev)
    pairs = float(N_POINTS) * N_POINTS

    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        evaluate_device(wall, x, x, eps_t, stride, f, u_re=u)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        evaluate_device(wall, x, x, eps_t, stride, f, u_re=u)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = 1e3 * wall_s / args.steps
    value = pairs / (wall_s / args.steps)
    from fluidlets_amd._lib import lib as _lib
    symmetric = os.environ.get("FL_EVAL_SYMMETRIC", "1") != "0" and _lib().fl_eval_sym_workspace_size(N_POINTS, 0, 0) > 0
    flops = sym_flops(N_POINTS, "wall") if symmetric else pairs * FLOPS_PER_PAIR["wall"]
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    traffic = profile_traffic("nbody_wall_c2_sym_traffic.txt" if symmetric else "nbody_wall_c2_traffic.txt")

    result = {
        "metric": "pair-evals/s (FP64)",
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform points, N(0,1) forces)",
        "config": {"workload": "N-body kernel-sum drop-in: WallStokeslet3D FP64, 262,144 sources = targets "
                               "uniform in [-4,4]^2x[0.05,4], eps=0.02, mu=1 (seed 2)",
                   "points": N_POINTS, "pairs_per_step": pairs, "kernel": "WallStokeslet3D", "eps": EPS},
        "roofline": {"bound": "valu", "pipe": "FP64 vector ALU (compute bound; HBM traffic is ~56 B per point)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": ("fl::nbody64_sym<WallStokeslet3D> (reciprocal: each cross-tile pair once, "
                                "applied both ways)") if symmetric else "fl::nbody64<WallStokeslet3D>",
                     "kernel_ms": kernel_ms, "flops_per_launch": flops,
                     "flops_per_pair": FLOPS_PER_PAIR["wall"],
                     "flops_per_cross_tile_pair": FLOPS_PER_PAIR["wall"] + FLOPS_TRANSPOSED_APPLY["wall"]
                     if symmetric else None},
    }

    if not args.no_table:
        table = {}
        ref = {}
        cases = [("wall", fl.WallStokeslet3D(), 1.0), ("stokeslet", fl.Stokeslet3D(), 1.0),
                 ("brinkman4", fl.Brinkmanlet3D(4.0), 1.0), ("stokeslet_mu1+0.5i", fl.Stokeslet3D(), 1 + 0.5j)]
        x_copy = x.clone()  # the same points as a separate buffer: the general (non-reciprocal) path
        for name, K, mu in cases:
            for prec in ["fp64", "fp64_general", "fp32c"]:
                src_t = x_copy if prec == "fp64_general" else x
                p_arg = "fp64" if prec == "fp64_general" else prec
                evaluate_device(K, x, src_t, eps_t, stride, f, mu=mu, precision=p_arg)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ur, ui = evaluate_device(K, x, src_t, eps_t, stride, f, mu=mu, precision=p_arg)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1)
                uu = ur if ui is None else torch.complex(ur, ui)
                row = {"pair_evals_per_s": pairs / (ms * 1e-3), "ms": ms}
                kfam = "wall" if name == "wall" else ("brinkman" if name.startswith("brinkman") else "stokeslet")
                fl_count = sym_flops(N_POINTS, kfam) if (symmetric and prec == "fp64") else pairs * FLOPS_PER_PAIR[kfam]
                row["tflops_fp64_convention"] = fl_count / (ms * 1e-3) / 1e12
                row["path"] = "reciprocal" if (symmetric and prec == "fp64") else "general"
                if prec == "fp64":
                    ref[name] = uu.clone()
                else:
                    row["rel_err_vs_fp64"] = (torch.linalg.norm(uu - ref[name]) / torch.linalg.norm(ref[name])).item()
                table[f"{name}/{prec}"] = row
        result["c2_table"] = table
        progress("kernel-sum table done")

    if not args.no_solve:
        result["icosphere_c1"] = c1_section(torch, fl)
        progress("icosphere solve done")
        result["wall_sphere_solve"] = c3_solves(torch, fl, D, tuple(args.solve_kernels.split(",")),
                                                args.solve_points, vendor=not args.no_vendor_lu,
                                                cpu_lapack=not args.no_cpu)
        progress("wall-sphere solves done")

    if not args.no_serve and args.small_solves_only:
        result["serving_mix"] = {"small_solves": small_solves_section()}
        progress("small solves done")
    elif not args.no_serve:
        result["serving_mix"] = serving_section(torch, args.serve_requests, args.serve_baseline_requests,
                                                args.serve_swimmer_steps, args.serve_nmax,
                                                cache_gib=args.serve_cache_gib, load=args.serve_load)
        progress("serving mix done")

    if not args.no_gmres:
        result["rod_carpet_gmres"] = gmres_section(torch, fl, D)
        progress("rod-carpet GMRES done")

    if not args.no_treecode and not args.small_solves_only:
        result["treecode_2m"] = treecode_section(torch, fl, D)
        progress(f"treecode at 2,097,152 points: {result['treecode_2m']['treecode_s']:.3f} s against "
                 f"{result['treecode_2m']['direct_s']:.3f} s direct, error {result['treecode_2m']['rel_err_vs_direct']:.1e}")

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(X, F, args.cpu_targets)
        result["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
    result["bench_wall_s"] = time.perf_counter() - t_bench
    print(json.dumps(result))


if __name__ == "__main__":
    main()
