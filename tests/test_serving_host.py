"""Host-side logic of the solve server (no GPU): footprints, the LRU of
resident factors, the request trace and the latency statistics."""
import numpy as np

from fluidlets_amd.serving import FactorCache, Request, Result, footprint_bytes, serving_mix, summarize


def test_footprint_scales_with_matrix_size():
    r = Request("resistance", {"points": np.zeros((1000, 3))})
    n = 3000
    assert footprint_bytes(r, "fp64") >= 16 * n * n
    assert footprint_bytes(r, "mixed") >= 12 * n * n
    assert footprint_bytes(r, "mixed") < footprint_bytes(r, "fp64")


def test_factor_cache_is_lru_and_bounded():
    c = FactorCache(100)
    c.put("a", ("A",), 40)
    c.put("b", ("B",), 40)
    assert c.get("a")[0] == "A"          # a is now most recent
    c.put("c", ("C",), 40)               # evicts b
    assert c.get("b") is None and c.get("c")[0] == "C" and c.bytes == 80
    c.put("huge", ("H",), 1000)          # larger than the cache: not stored
    assert c.get("huge") is None


def test_serving_mix_is_seeded_and_follows_the_spec():
    a = serving_mix(200, seed=5, swimmer_steps=2)
    b = serving_mix(200, seed=5, swimmer_steps=2)
    assert [r.kind for r in a] == [r.kind for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    frac = np.mean([r.kind == "swim_step" for r in a])
    assert 0.2 < frac < 0.4
    # swimmer requests are time steps of sessions of about swimmer_steps steps,
    # dealt round-robin, in order within each session
    steps = [r for r in a if r.kind == "swim_step"]
    sessions = sorted({r.payload["session"] for r in steps})
    assert len(sessions) == round(len(steps) / 2)
    for key in sessions:
        seq = [r.payload["seq"] for r in steps if r.payload["session"] == key]
        assert seq == list(range(len(seq))) and len(seq) in (2, 3)
    whole = serving_mix(200, seed=5, swimmer_steps=2, swimmer_mode="trajectory")
    assert [r.kind == "swimmer" for r in whole] == [r.kind == "swim_step" for r in a]
    for r in a:
        if r.kind == "resistance":
            X = r.payload["points"]
            assert 200 <= len(X) <= 6000
            if type(r.payload["kernel"]).__name__ == "WallStokeslet3D":
                assert X[:, 2].min() > r.payload["eps"] * 0.5       # every source above the wall
            np.testing.assert_array_equal(X, b[r.rid].payload["points"])


def test_summarize_percentiles():
    res = [Result(i, "resistance", None, 0.0, 0.0, float(i + 1)) for i in range(100)]
    res.append(Result(100, "resistance", None, 0.0, 0.0, 1.0, error="boom"))
    s = summarize(res)
    assert s["completed"] == 100 and s["failed"] == 1
    assert abs(s["latency_p50_s"] - 50.5) < 1e-9 and s["latency_p99_s"] > 99
