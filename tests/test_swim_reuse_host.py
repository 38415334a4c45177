"""Host logic of the swimmer's factor reuse (no GPU): the cache key of a body
shape and the frame's orthogonal polar factor."""
import numpy as np

from fluidlets_amd.cache import FactorCache
from fluidlets_amd.kernels import Brinkmanlet3D, Stokeslet3D, WallStokeslet3D
from fluidlets_amd.serving import Request, footprint_bytes
from fluidlets_amd.swimming import (SwimmerShape, _body_key, _frame_rotation, _rotation_invariant, biflagellate,
                                    reuse_cache_bytes)


def test_body_key_matches_one_beat_later_and_separates_phases():
    shape_at = biflagellate(n_body=100, n_flag=30)
    nn = np.arange(160)
    k0 = _body_key(Stokeslet3D(), shape_at(0.0), nn, np.array([0.02]), 1.0)
    # t = 1.0 is one period later: the same shape up to the rounding of sin/cos
    assert np.abs(shape_at(1.0).x - shape_at(0.0).x).max() < 1e-12
    assert _body_key(Stokeslet3D(), shape_at(1.0), nn, np.array([0.02]), 1.0) == k0
    assert _body_key(Stokeslet3D(), shape_at(0.02), nn, np.array([0.02]), 1.0) != k0
    assert _body_key(Stokeslet3D(), shape_at(0.0), nn, np.array([0.021]), 1.0) != k0
    assert _body_key(Stokeslet3D(), shape_at(0.0), nn, np.array([0.02]), 2.0) != k0
    assert _body_key(Brinkmanlet3D(1.0), shape_at(0.0), nn, np.array([0.02]), 1.0) != k0
    assert _body_key(Stokeslet3D(), shape_at(0.0), nn[::-1].copy(), np.array([0.02]), 1.0) != k0
    s = shape_at(0.0)
    moved = SwimmerShape(x=s.x + 1e-6, v=s.v, X=s.X + 1e-6)
    assert _body_key(Stokeslet3D(), moved, nn, np.array([0.02]), 1.0) != k0


def test_frame_rotation_is_the_polar_factor():
    rng = np.random.default_rng(2)
    Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    assert np.allclose(_frame_rotation(Q), Q, atol=1e-15)
    S = np.diag([1 + 1e-5, 1 + 1e-5, 1 + 2e-5])
    R = _frame_rotation(Q @ S)
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-15) and np.allclose(R, Q, atol=1e-12)
    # nearest orthogonal matrix to B in the Frobenius norm
    B = Q @ S
    for _ in range(20):
        P = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        assert np.linalg.norm(B - R) <= np.linalg.norm(B - P) + 1e-15


def test_reuse_applies_to_rotation_invariant_kernels_only():
    assert _rotation_invariant(Stokeslet3D()) and _rotation_invariant(Brinkmanlet3D(2.0))
    assert not _rotation_invariant(WallStokeslet3D())


def test_swimmer_footprint_counts_its_own_cache():
    req = Request("swimmer", {"n_points": 1800}, 0.0, 0)
    own = footprint_bytes(req, "mixed")
    shared = footprint_bytes(req, "mixed", shared_cache=True)
    assert own - shared == reuse_cache_bytes(1800) <= 16 << 30
    assert footprint_bytes(req, "fp64") == footprint_bytes(req, "fp64", shared_cache=True)


def test_factor_cache_replaces_a_key_without_leaking_bytes():
    c = FactorCache(100)
    c.put("a", (1,), 40)
    c.put("a", (2,), 50)
    assert c.bytes == 50 and c.items["a"][0] == 2
    c.put("b", (3,), 60)
    assert "a" not in c.items and c.bytes == 60


def test_factor_cache_makes_each_key_once_across_threads():
    """get_or_make: concurrent callers of one key wait for the first maker
    (one make, the rest hits); if that maker fails, its caller sees the
    error and one waiter makes the item instead."""
    import threading
    import time
    from fluidlets_amd.cache import FactorCache
    c = FactorCache(1000)
    calls = []

    def make():
        calls.append(1)
        time.sleep(0.05)
        return ("v",), 10

    out = []
    ths = [threading.Thread(target=lambda: out.append(c.get_or_make("k", make))) for _ in range(8)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert len(calls) == 1 and c.misses == 1 and c.hits == 7
    assert all(o[0] == "v" and o[-1] == 10 for o in out)

    c2 = FactorCache(1000)
    attempts = []

    def flaky():
        attempts.append(1)
        time.sleep(0.02)
        if len(attempts) == 1:
            raise RuntimeError("first attempt fails")
        return ("w",), 5

    res = []

    def run():
        try:
            res.append(c2.get_or_make("k", flaky)[0])
        except RuntimeError:
            res.append("err")

    ths = [threading.Thread(target=run) for _ in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert len(attempts) == 2 and sorted(res) == ["err", "w", "w", "w"]


def test_factor_cache_upgrade_replaces_an_item_once():
    """upgrade: the first caller replaces the stored item; a caller holding a
    stale item gets the current one; a caller arriving during an upgrade
    keeps its item without waiting."""
    import threading
    from fluidlets_amd.cache import FactorCache
    c = FactorCache(1000)
    made = []
    old = c.get_or_make("k", lambda: (("lu", None), 10), made)
    assert made == [True] and old[:2] == ("lu", None)
    new = c.upgrade("k", old, lambda: ((None, "inv"), 10))
    assert new[:2] == (None, "inv") and c.items["k"][:2] == (None, "inv") and c.bytes == 10
    assert c.upgrade("k", old, lambda: ((None, "again"), 10))[:2] == (None, "inv")   # stale: current item
    made2 = []
    assert c.get_or_make("k", lambda: (("x", None), 1), made2)[:2] == (None, "inv") and made2 == [False]

    c2 = FactorCache(1000)
    first = c2.get_or_make("k", lambda: (("lu", None), 10))
    started, release = threading.Event(), threading.Event()

    def slow():
        started.set()
        release.wait()
        return (None, "inv"), 10

    out = []
    th = threading.Thread(target=lambda: out.append(c2.upgrade("k", first, slow)))
    th.start()
    started.wait()
    assert c2.upgrade("k", first, lambda: ((None, "dup"), 10)) is first   # in flight: no wait
    release.set()
    th.join()
    assert out[0][:2] == (None, "inv") and c2.items["k"][:2] == (None, "inv")
