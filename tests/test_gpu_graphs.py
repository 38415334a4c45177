"""Small resistance solves captured as HIP graphs (fluidlets_amd.graphs) and
the fixed-step refinement they replay (fl_refine_steps).

Bars: every replay bit for bit equal to the same kernels launched directly,
with other work between replays; the padded solve within the refinement
tolerance of the oracle's dense FP64 solve of the unpadded matrix; resident
padded factors bit for bit the graph's result; fl_refine_steps bit for bit
fl_refine when fl_refine stops after the same number of corrections."""
import numpy as np
import pytest

import oracle as O
import fluidlets_amd as fl
from fluidlets_amd import graphs as G
from fluidlets_amd import linalg
from fluidlets_amd._lib import lib
from fluidlets_amd.core.assemble import assemble_padded, padded_order
from fluidlets_amd.core.solve import rigid_modes
from fluidlets_amd.geometry import ellipsoid, ellipsoid_area

pytestmark = pytest.mark.gpu


def body(N, seed, wall=True):
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.5, 2.0, 3)
    center = (0.0, 0.0, 2.5 + axes[2]) if wall else (0.0, 0.0, 0.0)
    eps = 0.5 * np.sqrt(ellipsoid_area(axes) / N)
    kernel = fl.WallStokeslet3D() if wall else fl.Brinkmanlet3D(float(rng.uniform(0, 10)))
    return kernel, ellipsoid(N, axes, center), center, eps


def test_refine_steps_is_refine_stopped_after_the_same_steps():
    import torch
    print("case: fl_refine_steps vs fl_refine n=1500", flush=True)
    kernel, X, c, eps = body(500, 3)
    K = O.assemble(kernel, X, X, eps)
    A64 = torch.from_numpy(K).cuda()
    F = linalg.getrf(A64.float(), refill=lambda A: A.copy_(A64))
    b = torch.from_numpy(rigid_modes(X, c)).cuda()
    x_ref, steps, res_ref = linalg.refine(A64, F, b, tol=1e-10)
    assert 1 <= steps <= 6
    x = torch.empty_like(b)
    resid = torch.zeros(8, dtype=torch.float64, device="cuda")
    ws = torch.empty(int(lib().fl_refine_workspace_size(len(K), 6)), dtype=torch.uint8, device="cuda")
    linalg.refine_steps(A64, F, b.contiguous(), x, resid, steps, ws)
    torch.cuda.synchronize()
    assert torch.equal(x, x_ref)
    assert np.array_equal(resid.cpu().numpy()[:6], res_ref)


def test_graph_replays_equal_direct_launches():
    """Three geometries of one bucket (different N, wall and Brinkman
    kernels), each replayed and launched directly with other device work in
    between, then replayed again: all equal bit for bit."""
    import torch
    print("case: graph replays vs direct launches, bucket of 1024 rows", flush=True)
    s = torch.cuda.Stream()
    g = G.SolveGraph(G.bucket_rows(300), s)
    assert g.n_pad == 1024 and g.graph is not None
    junk = torch.randn(1500, 1500, device="cuda", dtype=torch.float64)
    for N, seed, wall in [(300, 1, True), (331, 2, False), (341, 3, True)]:
        kernel, X, c, eps = body(N, seed, wall)
        modes = rigid_modes(X, c)
        x1, r1, _ = g.solve(kernel, X, eps, modes)
        (junk @ junk).sum().item()
        x2, r2, _ = g.solve(kernel, X, eps, modes, replay=False)
        linalg.getrf(junk.clone()).info
        x3, r3, _ = g.solve(kernel, X, eps, modes)
        assert np.array_equal(x1, x2) and np.array_equal(x1, x3), (N, seed)
        assert np.array_equal(r1, r2) and np.array_equal(r1, r3)
        assert np.all(r1 <= 1e-10)
    assert g.replays == 6


@pytest.mark.parametrize("N,wall", [(200, True), (451, False), (640, True)])
def test_padded_graph_solve_matches_the_dense_fp64_solve(N, wall):
    """The padded system's first 3N rows: residual of the unpadded FP64
    matrix within the tolerance and forward error within cond x tolerance of
    LAPACK's solve."""
    import torch
    print(f"case: padded graph solve N={N} wall={wall}", flush=True)
    kernel, X, c, eps = body(N, 10 + N, wall)
    K = O.assemble(kernel, X, X, eps)
    modes = rigid_modes(X, c)
    g = G.SolveGraph(G.bucket_rows(N), torch.cuda.Stream())
    x, res, _ = g.solve(kernel, X, eps, modes)
    ref = np.linalg.solve(K, modes)
    rel = np.linalg.norm(K @ x - modes, axis=0) / np.linalg.norm(modes, axis=0)
    assert np.all(rel <= 1.5e-10) and np.all(res <= 1e-10)
    cond = np.linalg.cond(K)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 2e-10 * cond


def test_graph_solve_with_per_point_blob_widths():
    """Blob widths given per point (staged and read with stride 1): the
    padded graph solve against the oracle's dense FP64 solve."""
    import torch
    print("case: graph solve with per-point eps", flush=True)
    kernel, X, c, eps = body(333, 77, True)
    eps_i = eps * np.random.default_rng(5).uniform(0.8, 1.2, len(X))
    K = O.assemble(kernel, X, X, eps_i)
    modes = rigid_modes(X, c)
    g = G.SolveGraph(G.bucket_rows(333), torch.cuda.Stream())
    x, res, _ = g.solve(kernel, X, eps_i, modes)
    rel = np.linalg.norm(K @ x - modes, axis=0) / np.linalg.norm(modes, axis=0)
    assert np.all(rel <= 1.5e-10) and np.all(res <= 1e-10)
    x_scalar, _, _ = g.solve(kernel, X, eps, modes)      # the same graph, one width again
    assert not np.allclose(x, x_scalar, rtol=1e-6)


def test_resident_padded_factors_reproduce_the_graph():
    import torch
    print("case: resident padded factors", flush=True)
    kernel, X, c, eps = body(410, 21, False)
    modes = rigid_modes(X, c)
    g = G.SolveGraph(G.bucket_rows(410), torch.cuda.Stream())
    x, res, F = g.solve(kernel, X, eps, modes)
    K64p, Fp = g.resident(F)
    other = body(400, 22, True)                    # the graph's buffers now hold another solve
    g.solve(other[0], other[1], other[3], rigid_modes(other[1], other[2]))
    y, res2 = G.solve_resident(K64p, Fp, modes)
    assert np.array_equal(x, y) and np.array_equal(res, res2)


def test_too_few_steps_fall_back_to_adaptive_refinement():
    """One correction in the graph leaves the residual above 1e-10: the
    solve continues with fl_refine from the start on the graph's factors."""
    import torch
    print("case: graph with one correction", flush=True)
    kernel, X, c, eps = body(300, 5, True)
    modes = rigid_modes(X, c)
    g = G.SolveGraph(G.bucket_rows(300), torch.cuda.Stream(), steps=1)
    x, res, F = g.solve(kernel, X, eps, modes)
    assert np.all(res <= 1e-10)
    xa, steps, _ = linalg.refine(g.K64, F, g.b, tol=1e-10)
    assert steps > 1
    assert np.array_equal(x, xa[:900].cpu().numpy())


def test_pool_captures_every_bucket_up_front_for_given_streams():
    import torch
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    pool = G.GraphPool(streams=[s1, s2])
    nb = G.bucket_rows(G.GRAPH_MAX_POINTS) // G.BUCKET_ROWS
    assert pool.captures == 2 * nb and not pool.lazy
    assert pool.nbytes() <= G.pool_bytes(2)
    g = pool.get(300, s2)
    assert g.n_pad == G.bucket_rows(300) and pool.get(300, s2) is g and pool.captures == 2 * nb
    with pytest.raises(KeyError):
        pool.get(300, torch.cuda.Stream())


def test_pool_rejects_large_solves_and_shares_graphs_per_stream():
    import torch
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    pool = G.GraphPool()
    with pytest.raises(ValueError):
        pool.get(G.GRAPH_MAX_POINTS + 1, s1)
    a = pool.get(200, s1)
    assert pool.get(210, s1) is a and pool.get(150, s1) is not a
    assert pool.get(200, s2) is not a and pool.captures == 3
    assert pool.nbytes() == 2 * G.graph_bytes(G.bucket_rows(200)) + G.graph_bytes(G.bucket_rows(150))


@pytest.mark.parametrize("N,multiple,per_point", [(331, 128, False), (700, 128, True), (64, 192, False)])
def test_assemble_padded_is_the_matrix_over_the_identity(N, multiple, per_point):
    """fl_assemble into the leading block and fl_pad_solve around it: the
    oracle's matrix, the identity below and to the right, the right-hand
    sides over zeros; the FP32 copy its rounding."""
    kernel, X, c, eps = body(N, 40 + N, N % 2 == 1)
    if per_point:
        eps = eps * np.random.default_rng(N).uniform(0.8, 1.2, N)
    modes = rigid_modes(X, c)
    K64, K32, b, n = assemble_padded(kernel, X, eps, modes, multiple=multiple)
    n_pad = padded_order(3 * N, multiple)
    assert n == 3 * N and tuple(K64.shape) == (n_pad, n_pad) and n_pad % multiple == 0
    P = np.eye(n_pad)
    P[:n, :n] = O.assemble(kernel, X, X, eps)
    got = K64.cpu().numpy()
    assert np.array_equal(got[n:], P[n:]) and np.array_equal(got[:n, n:], P[:n, n:])
    assert np.abs(got[:n, :n] - P[:n, :n]).max() <= 1e-13 * np.abs(P).max()
    assert np.array_equal(K32.cpu().numpy(), got.astype(np.float32))
    bb = b.cpu().numpy()
    assert np.array_equal(bb[:n], modes) and not bb[n:].any()
