"""Host side of the treecode (fluidlets_amd.treecode): the source tree and
the interaction lists, without a GPU.  The lists must cover every (target
leaf, source) pair exactly once, directly or through one cluster's proxies,
and take proxies only where the acceptance criterion holds."""
import numpy as np
import pytest

import fluidlets_amd as fl
from fluidlets_amd.treecode import Treecode, build_tree, chunk_table, interaction_lists


def slab(n, seed):
    return np.random.default_rng(seed).uniform([-4, -4, 0.05], [4, 4, 4], (n, 3))


@pytest.mark.parametrize("n,leaf", [(1, 4), (7, 4), (1000, 64), (5000, 300)])
def test_tree_partitions_the_points(n, leaf):
    X = slab(n, n)
    t = build_tree(X, leaf)
    assert np.array_equal(np.sort(t.order), np.arange(n))
    size = t.end - t.begin
    leaves = t.leaves
    assert size[leaves].max() <= leaf and size[leaves].sum() == n
    for k in np.flatnonzero(t.child[:, 0] >= 0):
        a, b = t.child[k]
        assert t.begin[a] == t.begin[k] and t.end[a] == t.begin[b] and t.end[b] == t.end[k]
        assert abs(int(size[a]) - int(size[b])) <= 1                       # median splits
    for k in range(t.begin.size):
        pts = X[t.order[t.begin[k]:t.end[k]]]
        if len(pts):
            assert np.all(pts >= t.lo[k]) and np.all(pts <= t.hi[k])     # tight boxes hold their points


@pytest.mark.parametrize("theta", [0.0, 0.5, 0.7, 0.9])
@pytest.mark.parametrize("same", [True, False])
def test_lists_cover_every_pair_once(theta, same):
    S = slab(6000, 1)
    T = S if same else np.random.default_rng(2).uniform([-6, -6, 0.05], [6, 6, 6], (4000, 3))
    st = build_tree(S, 200)
    tt = st if same else build_tree(T, 200)
    nq = 27
    (dB, dC), (aB, aC) = interaction_lists(tt, st, theta, nq)
    size = st.end - st.begin
    for lf in tt.leaves:
        cover = np.zeros(len(S), dtype=np.int64)
        for c in np.concatenate([dC[dB == lf], aC[aB == lf]]):
            cover[st.begin[c]:st.end[c]] += 1
        assert np.all(cover == 1), (lf, np.unique(cover))
    if theta == 0.0:
        assert aB.size == 0
    # proxies only for well-separated clusters larger than their proxy set
    d = np.linalg.norm(tt.centre[aB] - st.centre[aC], axis=1)
    assert np.all(tt.radius[aB] + st.radius[aC] < theta * d) and np.all(size[aC] > nq)


def test_larger_theta_sums_fewer_pairs_directly():
    S = slab(20000, 3)
    st = build_tree(S, 256)
    size = st.end - st.begin
    direct = []
    for th in (0.0, 0.3, 0.5, 0.7, 0.9):
        (dB, dC), _ = interaction_lists(st, st, th, 64)
        direct.append(int(np.sum(size[dC] * size[dB])))
    assert direct[0] == len(S) ** 2                     # theta = 0: the direct sum
    assert all(a > b for a, b in zip(direct, direct[1:]))


def test_treecode_needs_one_blob_width_and_a_valid_degree():
    X = slab(100, 4)
    with pytest.raises(ValueError, match="one blob width"):
        Treecode(fl.Stokeslet3D(), X, X, np.linspace(0.01, 0.02, 100))
    with pytest.raises(ValueError, match="degree"):
        Treecode(fl.Stokeslet3D(), X, X, 0.01, degree=13)
    with pytest.raises(ValueError, match="theta"):
        interaction_lists(build_tree(X, 10), build_tree(X, 10), 1.0, 8)


@pytest.mark.parametrize("target", [16384, 8])
def test_chunks_cover_each_cluster_in_order(target):
    """The moments kernel's chunks: consecutive per cluster, covering its
    range once, 4,096 sources (or a doubling of it) at most, and no more than
    ``target`` of them once the chunk size has grown."""
    rng = np.random.default_rng(7)
    begin = np.sort(rng.integers(0, 200000, 60))
    end = begin + rng.integers(730, 70000, 60)
    nch, first, ch = chunk_table(begin, end, target)
    ln = ch[:, 2] - ch[:, 1]
    last = first + nch - 1
    inner = np.setdiff1d(np.arange(ch.shape[0]), last)            # every chunk but a cluster's last
    size = int(ln[inner].max()) if inner.size else 1 << int(np.ceil(np.log2(max(ln.max(), 4096))))
    assert size >= 4096 and size & (size - 1) == 0 and np.all(ln[inner] == size) and np.all(ln <= size)
    assert ch.shape[0] == nch.sum() and (ch.shape[0] <= target or np.all(nch == 1))
    for c in range(begin.size):
        rows = ch[first[c]:first[c] + nch[c]]
        assert np.all(rows[:, 0] == c)
        assert rows[0, 1] == begin[c] and rows[-1, 2] == end[c]
        assert np.all(rows[1:, 1] == rows[:-1, 2]) and np.all(rows[:, 2] > rows[:, 1])


def test_no_clusters_no_chunks():
    nch, first, ch = chunk_table(np.zeros(0, np.int64), np.zeros(0, np.int64))
    assert nch.shape == first.shape == (0,) and ch.shape == (0, 3)
