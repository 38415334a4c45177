"""Grouped small solves (graphs.GroupSolveGraph over fl_pad_solve_group,
fl_sgetrf_async_group and fl_refine_steps_group) and the server that forms
the groups.

Bars: every member of a group bit for bit the same problem solved alone by
SolveGraph, whatever the group's size and the member's position, replayed
or launched directly; the members' residuals within the refinement
tolerance of the oracle's dense matrices; a grouping server's results bit for
bit those of a server that solves one request at a time."""
import numpy as np
import pytest

import oracle as O
import fluidlets_amd as fl
from fluidlets_amd import graphs as G
from fluidlets_amd.core.solve import rigid_modes
from fluidlets_amd.geometry import ellipsoid, ellipsoid_area
from fluidlets_amd.serving import Request, SolveServer, serving_mix

pytestmark = pytest.mark.gpu


def body(N, seed, wall=True, per_point=False):
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.5, 2.0, 3)
    center = (0.0, 0.0, 2.5 + axes[2]) if wall else (0.0, 0.0, 0.0)
    eps = 0.5 * np.sqrt(ellipsoid_area(axes) / N)
    if per_point:
        eps = eps * rng.uniform(0.8, 1.2, N)
    kernel = fl.WallStokeslet3D() if wall else fl.Brinkmanlet3D(float(rng.uniform(0, 10)))
    return kernel, ellipsoid(N, axes, center), center, eps


def problems(count, seed0=0, n_lo=300, n_hi=341):
    out = []
    rng = np.random.default_rng(seed0)
    for i in range(count):
        N = int(rng.integers(n_lo, n_hi + 1))
        kernel, X, c, eps = body(N, seed0 + i, wall=i % 2 == 0, per_point=i % 3 == 2)
        out.append((kernel, X, eps, rigid_modes(X, c)))
    return out


def test_group_members_equal_single_solves():
    """Five problems of the 1024-row bucket (wall and Brinkman kernels, one
    or per-point blob widths) in a group of eight slabs, then three of them
    in another order, then one alone, replayed and launched directly, with
    other device work in between: every member's x and residuals equal
    SolveGraph's bit for bit."""
    import torch
    print("case: group of 5, 3, 1 in the 1024-row bucket", flush=True)
    s = torch.cuda.Stream()
    grp = G.GroupSolveGraph(1024, s, members=8)
    assert sorted(grp.graphs) == list(range(1, 9))
    single = G.SolveGraph(1024, torch.cuda.Stream())
    probs = problems(5, seed0=11)
    ref = [single.solve(*p) for p in probs]
    junk = torch.randn(1500, 1500, device="cuda", dtype=torch.float64)
    for order, replay in (([0, 1, 2, 3, 4], True), ([4, 2, 0], True), ([3, 1], False), ([2], True)):
        out = grp.solve_many([probs[i] for i in order], replay=replay)
        (junk @ junk).sum().item()
        for i, (x, res, _) in zip(order, out):
            assert np.array_equal(x, ref[i][0]), (order, i)
            assert np.array_equal(res, ref[i][1]) and np.all(res <= 1e-10)
    assert grp.replays == 3


@pytest.mark.parametrize("n_pad,count", [(640, 8), (1920, 4)])
def test_group_members_solve_the_dense_systems(n_pad, count):
    """The smallest and the largest bucket: each member's first 3N rows
    against the oracle's dense FP64 matrix (relative residual 1.5e-10)."""
    import torch
    print(f"case: group of {count} in the {n_pad}-row bucket vs the oracle", flush=True)
    lo = (n_pad - G.BUCKET_ROWS) // 3 + 1
    probs = problems(count, seed0=n_pad, n_lo=lo, n_hi=n_pad // 3)
    grp = G.GroupSolveGraph(n_pad, torch.cuda.Stream(), members=count)
    for (kernel, X, eps, modes), (x, res, _) in zip(probs, grp.solve_many(probs)):
        assert G.bucket_rows(len(X)) == n_pad
        K = O.assemble(kernel, X, X, eps)
        rel = np.linalg.norm(K @ x - modes, axis=0) / np.linalg.norm(modes, axis=0)
        assert np.all(rel <= 1.5e-10) and np.all(res <= 1e-10)


def test_group_resident_factors_reproduce_the_member():
    import torch
    print("case: resident factors of a group member", flush=True)
    probs = problems(3, seed0=40)
    grp = G.GroupSolveGraph(1024, torch.cuda.Stream(), members=4)
    out = grp.solve_many(probs)
    K64p, Fp = grp.resident(1, out[1][2])
    grp.solve_many(problems(3, seed0=50))               # the slabs now hold other solves
    y, res = G.solve_resident(K64p, Fp, probs[1][3])
    assert np.array_equal(y, out[1][0]) and np.array_equal(res, out[1][1])


def test_group_rejects_what_does_not_fit():
    import torch
    grp = G.GroupSolveGraph(640, torch.cuda.Stream(), members=2, capture=False)
    with pytest.raises(ValueError):
        grp.solve_many(problems(3, seed0=1, n_lo=150, n_hi=200))     # more problems than slabs
    with pytest.raises(ValueError):
        grp.solve_many(problems(1, seed0=1, n_lo=300, n_hi=300))     # 900 rows > 640
    with pytest.raises(ValueError):
        grp.solve_many([])


def small_requests():
    reqs = serving_mix(24, seed=9, n_min=200, n_max=640, swimmer_frac=0.0, rate=1.0)
    assert all(r.kind == "resistance" for r in reqs)
    return reqs


def test_grouping_server_equals_one_at_a_time():
    """24 small requests queued at once on two slots with groups of up to
    eight: the groups form (fewer graph replays than requests) and every
    result equals the one-slot, one-request-at-a-time server's bit for bit;
    a repeated geometry is then served from its resident factors."""
    print("case: grouping server vs one at a time", flush=True)
    reqs = small_requests()
    with SolveServer(slots=1, precision="mixed") as one:
        ref = one.serve(reqs, realtime=False)
    with SolveServer(slots=2, precision="mixed", group=8, cache_bytes=4 << 30) as srv:
        out = srv.serve(reqs, realtime=False)
        assert all(r.error is None for r in out), [r.error for r in out if r.error]
        assert srv.graphs.replays < len(reqs)
        for a, b in zip(ref, out):
            assert a.error is None and np.array_equal(a.value, b.value), a.rid
        again = srv.serve([Request(r.kind, r.payload, 0.0, r.rid) for r in reqs[:5]], realtime=False)
        assert srv.cache.hits >= 5
        for a, b in zip(ref[:5], again):
            assert np.array_equal(a.value, b.value)
