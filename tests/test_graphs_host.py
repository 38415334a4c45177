"""The padding behind the captured small solves, checked with LAPACK on the
CPU: partial pivoting on [[K, 0], [0, I]] never takes a padding row in the
first n columns, the factors are K's next to I, and the padded solution is
K's solution over zeros.  Random Gaussian K, so that every column pivots."""
import numpy as np
import pytest
import scipy.linalg as sl

from fluidlets_amd import graphs as G


def test_padded_order():
    from fluidlets_amd.core.assemble import PAD_MULTIPLE, padded_order
    assert PAD_MULTIPLE == 128
    assert [padded_order(n) for n in (1, 128, 129, 1926, 5403, 24576)] == [128, 128, 256, 2048, 5504, 24576]
    assert padded_order(300, 192) == 384


def test_bucket_rows():
    assert [G.bucket_rows(n) for n in (1, 64, 65, 128, 300, 640)] == [128, 256, 256, 384, 1024, 1920]
    assert G.bucket_rows(G.GRAPH_MAX_POINTS) == 3 * G.GRAPH_MAX_POINTS


@pytest.mark.parametrize("n,n_pad", [(7, 192), (300, 384), (1157, 1920)])
def test_identity_padding_keeps_the_factors_and_the_solution(n, n_pad):
    rng = np.random.default_rng(n)
    K = rng.standard_normal((n, n))
    P = np.eye(n_pad)
    P[:n, :n] = K
    P[:n, n:] = 0.0
    lu, piv = sl.lu_factor(K)
    lu_p, piv_p = sl.lu_factor(P)
    assert np.array_equal(piv_p[:n], piv)                  # the same interchanges, none with padding rows
    assert np.array_equal(piv_p[n:], np.arange(n, n_pad))  # the padding is not interchanged
    assert np.allclose(lu_p[:n, :n], lu, rtol=0, atol=1e-12 * np.abs(lu).max())
    assert np.array_equal(lu_p[n:, :], P[n:, :]) and np.array_equal(lu_p[:n, n:], np.zeros((n, n_pad - n)))
    b = rng.standard_normal((n, 6))
    bp = np.zeros((n_pad, 6))
    bp[:n] = b
    x = sl.lu_solve((lu_p, piv_p), bp)
    assert np.array_equal(x[n:], np.zeros((n_pad - n, 6)))
    ref = np.linalg.solve(K, b)
    assert np.linalg.norm(x[:n] - ref) <= 1e-13 * np.linalg.cond(K) * np.linalg.norm(ref)


def test_lazy_pool_captures_only_in_its_creating_thread():
    """A capture that overlaps other threads' runtime calls can end the
    process, so a pool without streams refuses to capture from another
    thread (checked before any GPU call)."""
    import threading

    class FakeStream:
        cuda_stream = 12345

    pool = G.GraphPool()
    seen = []

    def other():
        try:
            pool.get(100, FakeStream())
        except Exception as exc:       # noqa: BLE001 - the exception type is the check
            seen.append(exc)

    th = threading.Thread(target=other)
    th.start()
    th.join()
    assert len(seen) == 1 and isinstance(seen[0], RuntimeError) and "creat" in str(seen[0])
    assert pool.captures == 0
