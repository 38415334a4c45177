"""Parity of the HIP kernel sums (fl_eval through the C ABI) with the oracle.

FP64 bar: normwise relative error <= 1e-12 against the C restatement of the
reference loop (and the legacy golden data at the reference's own
tolerances).  FP32C bar: normwise relative error <= 2e-6 against FP64.
"""
import os

import numpy as np
import pytest

import oracle as O
import fluidlets_amd as fl
from conftest import golden
from helpers import rel_err

pytestmark = pytest.mark.gpu

KERNELS = {
    "stokeslet": fl.Stokeslet3D(),
    "wall": fl.WallStokeslet3D(),
    "brinkman:0": fl.Brinkmanlet3D(0.0),
    "brinkman:0.5": fl.Brinkmanlet3D(0.5),
    "brinkman:4.0": fl.Brinkmanlet3D(4.0),
    "brinkman:20.0": fl.Brinkmanlet3D(20.0),
}


def cloud(rng, n, lo=(-1, -1, 0.05), hi=(1, 1, 1.5)):
    return rng.uniform(lo, hi, (n, 3))


def close(u, ref, tol=1e-12):
    assert u.shape == ref.shape and u.dtype == ref.dtype
    e = rel_err(u, ref)
    assert e <= tol, e
    m = np.abs(u - ref).max() / np.abs(ref).max()
    assert m <= 10 * tol, m


def test_legacy_matlab_fixtures(kernels_mat):
    """tests/legacy/test_phase1_port.py:33-44 at the reference's tolerances."""
    print("case: legacy kernels.mat", flush=True)
    d = kernels_mat
    u = fl.evaluate(fl.Stokeslet3D(), d["targets"], d["sources"], d["strengths"], d["deltas"])
    assert np.allclose(u, d["V_stokeslet"], rtol=1e-13, atol=1e-15)
    u = fl.evaluate(fl.WallStokeslet3D(), d["targets"], d["sources"], d["strengths"], d["deltas"])
    assert np.allclose(u, d["V_stokeslet"] + d["V_image"], rtol=1e-13, atol=1e-15)
    u = fl.evaluate(fl.WallStokeslet3D(), d["wall"], d["sources"], d["strengths"], d["deltas"])
    assert np.allclose(u, d["V_wall_total"], atol=1e-15)
    u = fl.stokeslet3d_velocity(d["targets"], d["sources"], d["strengths"], d["deltas"])
    assert np.allclose(u, d["V_stokeslet"], rtol=1e-13, atol=1e-15)


def test_legacy_tower_operator_and_grid(towers_mat):
    """tests/legacy/test_phase1_port.py:54-59 and 77-87."""
    print("case: legacy towers", flush=True)
    d = towers_mat
    W = fl.WallStokeslet3D()
    r = np.linalg.norm(fl.evaluate(W, d["X"], d["X"], d["F"], 0.5) - d["U"]) / np.linalg.norm(d["U"])
    assert abs(r - d["resid"]) < 1e-12
    xs, ys, zs = np.linspace(-1, 8, 32), np.linspace(-5, 25, 32), np.linspace(0, 20, 32)
    Z, Y, Xg = np.meshgrid(zs, ys, xs, indexing="ij")
    grid = np.c_[Xg.ravel(), Y.ravel(), Z.ravel()]
    V = fl.evaluate(W, grid, d["X"], d["F"], 0.5)
    V[:, 0] += 0.1 * grid[:, 2]
    stored = golden("legacy_grid.npz")["velocity"].astype(np.float64).reshape(-1, 3)
    assert np.linalg.norm(V - stored) / np.linalg.norm(stored) < 1e-6
    close(V - np.c_[0.1 * grid[:, 2], 0 * grid[:, :2]], O.evaluate("wall", grid, d["X"], d["F"], 0.5))


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("M,N", [(1, 1), (257, 513), (1000, 3000), (96, 40000)])
def test_fp64_matches_oracle(name, M, N):
    print(f"case: fp64 {name} M={M} N={N}", flush=True)
    rng = np.random.default_rng(M * 7 + N)
    X, Y = cloud(rng, M), cloud(rng, N)
    F = rng.standard_normal((N, 3))
    eps = rng.uniform(0.02, 0.2, N)
    close(fl.evaluate(KERNELS[name], X, Y, F, eps), O.evaluate(name, X, Y, F, eps))
    close(fl.evaluate(KERNELS[name], X, Y, F, 0.05), O.evaluate(name, X, Y, F, 0.05))


@pytest.mark.parametrize("name", ["stokeslet", "wall", "brinkman:4.0"])
def test_complex_viscosity_and_strengths(name):
    """tests/test_kernels.py:45-52: u(mu) = u(1)/mu, complex128 results."""
    print(f"case: complex {name}", flush=True)
    rng = np.random.default_rng(3)
    X, Y = cloud(rng, 300), cloud(rng, 700)
    F = rng.standard_normal((700, 3))
    Fc = F + 1j * rng.standard_normal((700, 3))
    for mu in [2.0 - 0.5j, 1 + 0.5j, 0.7]:
        u = fl.evaluate(KERNELS[name], X, Y, F, 0.1, mu=mu)
        assert u.dtype == (np.complex128 if isinstance(mu, complex) else np.float64)
        close(u, O.evaluate(name, X, Y, F, 0.1, mu=mu))
        uc = fl.evaluate(KERNELS[name], X, Y, Fc, 0.1, mu=mu)
        assert uc.dtype == np.complex128
        close(uc, O.evaluate(name, X, Y, Fc, 0.1, mu=mu))


def test_empty_and_degenerate_inputs():
    K = fl.WallStokeslet3D()
    assert fl.evaluate(K, np.zeros((0, 3)), np.ones((4, 3)), np.ones((4, 3)), 0.1).shape == (0, 3)
    u = fl.evaluate(K, np.ones((5, 3)), np.zeros((0, 3)), np.zeros((0, 3)), 0.1)
    assert u.shape == (5, 3) and not u.any()
    # coincident target and source (the r = 0 self term is finite)
    x = np.array([[0.2, -0.1, 0.5]])
    close(fl.evaluate(K, x, x, [[1.0, 2.0, 3.0]], 0.1), O.evaluate("wall", x, x, [[1.0, 2.0, 3.0]], 0.1))
    with pytest.raises(ValueError):
        fl.evaluate(K, np.ones((3, 2)), np.ones((3, 3)), np.ones((3, 3)), 0.1)
    with pytest.raises(ValueError):
        fl.Brinkmanlet3D(-1.0)


def test_wall_velocity_vanishes_on_the_wall():
    """tests/test_kernels.py:39-42."""
    rng = np.random.default_rng(7)
    wall = np.c_[rng.uniform(-2, 2, (50, 2)), np.zeros(50)]
    src = np.c_[rng.uniform(-1, 1, (10, 2)), rng.uniform(0.2, 1.5, 10)]
    u = fl.evaluate(fl.WallStokeslet3D(), wall, src, rng.normal(size=(10, 3)), 0.1)
    assert np.abs(u).max() < 1e-14


def test_brinkman_alpha_zero_is_stokeslet_and_series_branch():
    """tests/test_brinkmanlet.py:14-18; targets near sources exercise z < 0.5."""
    rng = np.random.default_rng(11)
    X, X0, F = rng.normal(size=(25, 3)), rng.normal(size=(15, 3)), rng.normal(size=(15, 3))
    assert np.allclose(fl.evaluate(fl.Brinkmanlet3D(0.0), X, X0, F, 0.1),
                       fl.evaluate(fl.Stokeslet3D(), X, X0, F, 0.1), rtol=1e-13, atol=1e-15)
    src = rng.normal(size=(400, 3))
    tgt = src + 0.01 * rng.normal(size=src.shape)
    for a in [1.0, 4.0, 10.0]:
        close(fl.evaluate(fl.Brinkmanlet3D(a), tgt, src, F[:1].repeat(400, 0), 0.05),
              O.evaluate(f"brinkman:{a}", tgt, src, F[:1].repeat(400, 0), 0.05))


@pytest.mark.parametrize("tpl", ["1", "2"])
def test_register_blocking_paths_agree(tpl, monkeypatch):
    """Both targets-per-lane variants (FL_NBODY_TPL) give the oracle's answer."""
    monkeypatch.setenv("FL_NBODY_TPL", tpl)
    rng = np.random.default_rng(5)
    X, Y = cloud(rng, 5000), cloud(rng, 6000)
    F = rng.standard_normal((6000, 3))
    for name in ["stokeslet", "wall", "brinkman:4.0"]:
        close(fl.evaluate(KERNELS[name], X, Y, F, 0.03), O.evaluate(name, X, Y, F, 0.03))


@pytest.mark.parametrize("name", ["stokeslet", "wall", "brinkman:4.0"])
def test_fp32c_error_against_fp64(name):
    print(f"case: fp32c {name}", flush=True)
    rng = np.random.default_rng(2)
    X = rng.uniform([-4, -4, 0.05], [4, 4, 4], (4096, 3))
    F = rng.standard_normal((4096, 3))
    u64 = fl.evaluate(KERNELS[name], X, X, F, 0.02)
    u32 = fl.evaluate(KERNELS[name], X, X, F, 0.02, precision="fp32c")
    assert rel_err(u32, u64) < 2e-6


def test_full_size_c2_wall_subset_and_invariants():
    """The 262,144-point workload: the first 512 targets against the oracle
    over all sources, plus size-independent checks (linearity, zero velocity
    at targets on the wall)."""
    import torch
    from fluidlets_amd import _device as D
    from fluidlets_amd.core.matvec import evaluate_device
    print("case: C2 full size wall", flush=True)
    rng = np.random.default_rng(2)
    n = 262144
    X = rng.uniform([-4, -4, 0.05], [4, 4, 4], (n, 3))
    F = rng.standard_normal((n, 3))
    dev = D.device()
    x = D.soa(X, dev)
    f = D.soa(F, dev)
    eps_t, st = D.eps_arg(0.02, n, dev)
    K = fl.WallStokeslet3D()
    u, _ = evaluate_device(K, x, x, eps_t, st, f)
    torch.cuda.synchronize()
    u_host = D.to_host_aos(u)
    # first, a middle and the last (ragged-free) tile: row partials only, both, column partials only
    for sl in [slice(0, 512), slice(131000, 131200), slice(n - 300, n)]:
        close(u_host[sl], O.evaluate("wall", X[sl], X, F, 0.02))
    # linearity: u(2f1 - 3f2) = 2u(f1) - 3u(f2)
    g = D.soa(rng.standard_normal((n, 3)), dev)
    ug, _ = evaluate_device(K, x, x, eps_t, st, g)
    ul, _ = evaluate_device(K, x, x, eps_t, st, 2 * f - 3 * g)
    assert (torch.linalg.norm(ul - (2 * u - 3 * ug)) / torch.linalg.norm(ul)).item() < 1e-13
    # wall targets see no flow
    W = np.c_[rng.uniform(-4, 4, (1000, 2)), np.zeros(1000)]
    uw, _ = evaluate_device(K, D.soa(W, dev), x, eps_t, st, f)
    assert uw.abs().max().item() < 1e-11 * u.abs().max().item()


@pytest.mark.parametrize("name", ["stokeslet", "wall", "brinkman:4.0", "brinkman:20.0"])
def test_symmetric_path_matches_general_path_and_oracle(name):
    """Targets = sources with one blob width takes the reciprocal (each pair
    once, applied both ways) evaluation; it must agree with the general path
    (same points passed as a separate copy) and with the oracle on target
    subsets in the first, a middle and the last, ragged tile.  N = 30,011 is
    not a multiple of the 512-point tile."""
    import torch
    from fluidlets_amd import _device as D
    from fluidlets_amd._lib import lib
    from fluidlets_amd.core.matvec import evaluate_device
    print(f"case: symmetric path {name}", flush=True)
    n = 30011
    assert lib().fl_eval_sym_workspace_size(n, 0, 0) > 0
    rng = np.random.default_rng(11)
    X = rng.uniform([-2, -2, 0.05], [2, 2, 2], (n, 3))
    X[7] = X[6]                                  # a coincident pair
    F = rng.standard_normal((n, 3))
    dev = D.device()
    x, f = D.soa(X, dev), D.soa(F, dev)
    eps_t, st = D.eps_arg(0.03, n, dev)
    K = KERNELS[name]
    us, _ = evaluate_device(K, x, x, eps_t, st, f)
    ug, _ = evaluate_device(K, x, x.clone(), eps_t, st, f)
    us2, _ = evaluate_device(K, x, x, eps_t, st, f)
    torch.cuda.synchronize()
    assert torch.equal(us, us2)                  # deterministic
    assert (torch.linalg.norm(us - ug) / torch.linalg.norm(ug)).item() < 1e-13
    uh = D.to_host_aos(us)
    for sl in [slice(0, 300), slice(15000, 15300), slice(n - 300, n)]:
        close(uh[sl], O.evaluate(name, X[sl], X, F, 0.03))
    # complex viscosity through the same path
    uc = fl.evaluate(K, X, X, F, 0.03, mu=1 + 0.5j)
    assert np.allclose(uc, uh / (1 + 0.5j), rtol=1e-13, atol=1e-15 * np.abs(uh).max())


def test_symmetric_path_in_gmres():
    """Matrix-free GMRES on 23,552 wall points uses the reciprocal matvec: the
    solution meets the residual bar and agrees with the general-path solve."""
    import torch
    from fluidlets_amd import _device as D
    from fluidlets_amd.core.gmres import gmres_device
    from fluidlets_amd.geometry import rod_carpet
    print("case: symmetric matvec in GMRES", flush=True)
    X = rod_carpet(n_side=16, per_rod=92)
    n = len(X)
    assert n == 23552
    dev = D.device()
    x = D.soa(X, dev)
    eps_t, st = D.eps_arg(0.01, n, dev)
    U = np.zeros((n, 3))
    U[:, 0] = -X[:, 2]
    U_d = D.soa(U, dev)
    K = fl.WallStokeslet3D()
    f1, it1, r1 = gmres_device(K, x, eps_t, st, U_d, rtol=1e-9, restart=50, maxiter=60)
    os.environ["FL_EVAL_SYMMETRIC"] = "0"
    try:
        f2, it2, r2 = gmres_device(K, x, eps_t, st, U_d, rtol=1e-9, restart=50, maxiter=60)
    finally:
        del os.environ["FL_EVAL_SYMMETRIC"]
    torch.cuda.synchronize()
    assert r1 <= 1e-9 and r2 <= 1e-9
    assert (torch.linalg.norm(f1 - f2) / torch.linalg.norm(f2)).item() < 1e-7
    u = fl.evaluate(K, X, X, D.to_host_aos(f1), 0.01)
    assert np.linalg.norm(u - U) / np.linalg.norm(U) < 2e-9
