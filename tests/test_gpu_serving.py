"""The solve server on the GPU: requests overlap on several streams and each
result is bit-for-bit the result of the same request served alone."""
import numpy as np
import pytest

import fluidlets_amd as fl
from fluidlets_amd.serving import Request, SolveServer, serving_mix, summarize

pytestmark = pytest.mark.gpu


def small_mix():
    reqs = serving_mix(10, seed=7, n_min=200, n_max=900, swimmer_steps=2, swimmer_body=150, swimmer_flag=40,
                       rate=200.0, swimmer_mode="trajectory")
    assert any(r.kind == "swimmer" for r in reqs) and any(r.kind == "resistance" for r in reqs)
    return reqs


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_results_alone_equal_results_in_the_mix(precision):
    print(f"case: serving alone vs mix {precision}", flush=True)
    reqs = small_mix()
    srv = SolveServer(slots=4, precision=precision)
    mixed = srv.serve(reqs, realtime=True)
    assert all(r.error is None for r in mixed), [r.error for r in mixed]
    if precision == "mixed":       # the small solves replayed captured graphs
        small = [r for r in reqs if r.kind == "resistance" and len(r.payload["points"]) <= 640]
        assert small and srv.graphs.replays == len(small)
    alone = SolveServer(slots=1, precision=precision)
    for r in reqs:
        ref = alone.serve([Request(r.kind, r.payload, 0.0, r.rid)], realtime=False)[0]
        assert ref.error is None
        assert np.array_equal(ref.value, mixed[r.rid].value), r.rid
    s = summarize(mixed)
    assert s["completed"] == len(reqs) and s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_graph_and_direct_small_solves_agree():
    """Small solves through the captured graphs (padded, fixed corrections)
    and through the unpadded adaptive path: the same grand resistance to
    the refinement tolerance."""
    reqs = [r for r in small_mix() if r.kind == "resistance"]
    reqs = [r for r in reqs if len(r.payload["points"]) <= 640][:3]
    assert reqs
    a = SolveServer(slots=2, precision="mixed").serve(reqs, realtime=False)
    b = SolveServer(slots=2, precision="mixed", graphs=False).serve(reqs, realtime=False)
    for x, y in zip(a, b):
        assert x.error is None and y.error is None
        assert np.allclose(x.value, y.value, rtol=0, atol=1e-9 * np.abs(y.value).max())


def test_padded_direct_solves_match_the_library():
    """Requests above the graphs' range run identity-padded to a multiple of
    128 rows: the same grand resistance as the library's unpadded FP64
    solve, to the refinement tolerance (mixed) and to rounding (FP64)."""
    reqs = [r for r in small_mix() if r.kind == "resistance" and len(r.payload["points"]) > 640][:2]
    assert len(reqs) == 2 and all(3 * len(r.payload["points"]) % 128 for r in reqs)
    for precision, tol in (("mixed", 1e-9), ("fp64", 1e-12)):
        out = SolveServer(slots=2, precision=precision).serve(reqs, realtime=False)
        for r, o in zip(reqs, out):
            assert o.error is None, o.error
            p = r.payload
            R, _, _ = fl.grand_resistance(p["kernel"], p["points"], p["center"], p["eps"])
            assert np.allclose(o.value, R, rtol=0, atol=tol * np.abs(R).max()), precision


def test_grand_resistance_matches_the_library_solve():
    reqs = [r for r in small_mix() if r.kind == "resistance"][:2]
    out = SolveServer(slots=2, precision="fp64").serve(reqs, realtime=False)
    for r, o in zip(reqs, out):
        p = r.payload
        R, _, _ = fl.grand_resistance(p["kernel"], p["points"], p["center"], p["eps"])
        assert np.allclose(o.value, R, rtol=1e-12, atol=1e-12 * np.abs(R).max())


def test_resident_factors_serve_repeated_geometry():
    reqs = [r for r in small_mix() if r.kind == "resistance"][:1]
    again = [Request(reqs[0].kind, reqs[0].payload, 0.0, 1)]
    srv = SolveServer(slots=1, precision="mixed", cache_bytes=2 << 30)
    a = srv.serve([Request(reqs[0].kind, reqs[0].payload, 0.0, 0)], realtime=False)[0]
    b = srv.serve(again, realtime=False)[0]
    assert srv.cache.hits == 1 and srv.cache.misses == 1
    assert np.array_equal(a.value, b.value)


def test_admission_and_failures_do_not_stop_the_server():
    reqs = [r for r in small_mix() if r.kind == "resistance"][:3]
    big = Request("resistance", dict(reqs[0].payload, points=np.zeros((40000, 3))), 0.0, 10)
    bad = Request("resistance", dict(reqs[1].payload, kernel="not a kernel"), 0.0, 11)
    out = SolveServer(slots=2, precision="mixed", budget_bytes=8 << 30).serve(reqs + [big, bad], realtime=False)
    by = {r.rid: r for r in out}
    assert by[10].error.startswith("rejected") and by[11].error is not None
    assert all(by[r.rid].error is None for r in reqs)


def test_concurrent_swimmers_share_each_shape_once():
    """Two identical swimmers served together over the server's cache: every
    body shape is made once (the other swimmer waits for it), and the two
    trajectories agree to the refinement tolerance."""
    from fluidlets_amd.swimming import biflagellate
    print("case: serving two swimmers, shared cache", flush=True)
    shape_at = biflagellate(n_body=150, n_flag=40)
    pay = {"shape_at": shape_at, "dt": 1.0 / 8, "steps": 12, "eps": 0.02, "n_points": 230}
    srv = SolveServer(slots=2, precision="mixed", cache_bytes=2 << 30)
    out = srv.serve([Request("swimmer", pay, 0.0, 0), Request("swimmer", pay, 0.0, 1)], realtime=False)
    assert all(r.error is None for r in out), [r.error for r in out]
    assert srv.cache.misses == 16                # 16 stage shapes per beat (dt = T/8)
    a, b = out[0].value, out[1].value
    assert np.allclose(a, b, rtol=0, atol=1e-11 * np.abs(a).max())
