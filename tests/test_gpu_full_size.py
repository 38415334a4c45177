"""The solver workloads at their full sizes, through properties that do not
depend on an oracle solve of that size:

- the wall-sphere resistance problem (8,192 Fibonacci points at height 1.5,
  n = 24,576, six rigid-body right-hand sides): the mixed-precision solve
  (FP32 LU with split-bf16 trailing products, FP64 refinement) within 1e-10
  of the FP64 LU, whose own residual is at rounding level; the grand
  resistance matrices agree to 1e-12, and the FP64 one has the structure of
  a sphere above a wall (positive translational and rotational diagonals,
  vertical drag above lateral drag);
- the rod carpet (1,024 rods of 128 points, 131,072 points, GMRES(50) to
  1e-8): block-Jacobi preconditioned GMRES reaches the tolerance in fewer
  iterations than plain GMRES, and the two force fields agree to the
  tolerance's reach (1e-6)."""
import numpy as np
import pytest

import fluidlets_amd as fl
from fluidlets_amd.geometry import fibonacci_sphere, rod_carpet
from helpers import rel_err

pytestmark = pytest.mark.gpu


def test_full_size_wall_sphere_mixed_equals_fp64():
    import torch
    print("case: wall sphere N=8192 mixed vs fp64", flush=True)
    N = 8192
    X = fibonacci_sphere(N, 1.0, (0.0, 0.0, 1.5))
    eps = 0.5 * np.sqrt(4 * np.pi / N)
    c = (0.0, 0.0, 1.5)
    R64, f64, _ = fl.grand_resistance(fl.WallStokeslet3D(), X, c, eps, precision="fp64")
    torch.cuda.empty_cache()
    Rm, fm, info = fl.grand_resistance(fl.WallStokeslet3D(), X, c, eps, precision="mixed", tol=1e-12)
    torch.cuda.empty_cache()
    assert rel_err(fm, f64) < 1e-10
    assert np.abs(Rm - R64).max() < 1e-12 * np.abs(R64).max()
    assert 2 <= info["refinement_steps"] <= 5
    # structure of a sphere above a wall
    d = np.diag(R64)
    assert np.all(d > 0)
    assert d[2] > d[0] and abs(d[0] - d[1]) < 1e-6 * d[0]
    # the FP64 residual of one mode at rounding level (matrix-free, by the kernel sum)
    U = np.zeros((N, 3))
    U[:, 2] = 1.0
    u = fl.evaluate(fl.WallStokeslet3D(), X, X, f64[2], eps)
    assert rel_err(u, U) < 1e-12


def test_full_size_rod_carpet_block_jacobi_gmres():
    from fluidlets_amd import _device as D
    from fluidlets_amd.core.gmres import gmres_device
    import torch
    print("case: rod carpet 131072 points GMRES", flush=True)
    X = rod_carpet(n_side=32, per_rod=128)
    N = len(X)
    assert N == 131072
    dev = D.device()
    x_soa = D.soa(X, dev)
    eps_t, stride = D.eps_arg(0.01, N, dev)
    U = np.zeros((N, 3))
    U[:, 0] = -X[:, 2]
    U_soa = D.soa(U, dev)
    f0, it0, r0 = gmres_device(fl.WallStokeslet3D(), x_soa, eps_t, stride, U_soa, rtol=1e-8, restart=50, maxiter=40)
    f1, it1, r1 = gmres_device(fl.WallStokeslet3D(), x_soa, eps_t, stride, U_soa, rtol=1e-8, restart=50, maxiter=40,
                               block_points=128)
    assert r0 <= 1e-8 and r1 <= 1e-8
    assert it1 < it0
    assert (torch.linalg.norm(f1 - f0) / torch.linalg.norm(f0)).item() < 1e-6


def test_full_size_general_fp64_lu_residual():
    """The general FP64 LU (row interchanges, no symmetric shortcut) of the
    n = 24,576 Brinkmanlet sphere matrix solves the six rigid-body modes to
    a relative residual below 1e-11, checked by a vendor FP64 product with
    the assembled matrix.  Its panel products run at full size here (some
    in place on their B operand, with several hundred workgroups), which the
    smaller solver tests do not reach."""
    import torch
    from fluidlets_amd import _device as D, linalg
    from fluidlets_amd.core.assemble import assemble_device
    from fluidlets_amd.core.solve import rigid_modes
    print("case: general FP64 LU n=24576 residual", flush=True)
    N = 8192
    X = fibonacci_sphere(N, 1.0, (0.0, 0.0, 0.0))
    dev = D.device()
    x = D.soa(X, dev)
    eps_t, stride = D.eps_arg(0.5 * np.sqrt(4 * np.pi / N), N, dev)
    rhs = torch.from_numpy(rigid_modes(X, (0.0, 0.0, 0.0))).to(dev)
    K64, K32 = assemble_device(fl.Brinkmanlet3D(5.0), x, x, eps_t, stride, 1.0, fp64=True, fp32=True)
    del K32
    A = K64.clone()
    F = linalg.getrf(A, refill=lambda Y: Y.copy_(K64))
    sol = rhs.clone()
    linalg.getrs(F, sol)
    res = (rhs - K64 @ sol).norm(dim=0) / rhs.norm(dim=0)
    print("relative residuals", res.cpu().numpy(), flush=True)
    assert float(res.max()) < 1e-11
    del A, F, K64
    torch.cuda.empty_cache()
