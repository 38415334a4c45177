"""Block-Jacobi preconditioning of GMRES (fl_assemble_blocks, fl_block_lu,
fl_block_apply, fl_gmres_bj) against LAPACK (scipy), per-block dense solves
and the oracle's dense resistance solve.  Bars: pivots equal to LAPACK's and
factors within 1e-12; applies backward stable and within cond(A) x 1e-15 of
numpy's solve; the
preconditioned GMRES solution within the GMRES tolerance of the dense solve,
in fewer iterations than plain GMRES."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
import fluidlets_amd as fl
from fluidlets_amd._lib import check, kernel_struct, lib
from fluidlets_amd.geometry import rod_carpet
from helpers import rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def block_lu(blocks):
    """Factor a (nblk, m, m) host array on the device."""
    import torch
    nblk, m, _ = blocks.shape
    KB = dev(blocks.copy())
    ipiv = torch.empty((nblk, m), dtype=torch.int32, device="cuda")
    perm = torch.empty_like(ipiv)
    info = torch.empty((nblk,), dtype=torch.int32, device="cuda")
    check(lib().fl_block_lu(m, nblk, KB.data_ptr(), ipiv.data_ptr(), perm.data_ptr(), info.data_ptr(), None),
          "fl_block_lu")
    torch.cuda.synchronize()
    return KB, ipiv, perm, info


@pytest.mark.parametrize("P", [1, 11, 32, 128, 170])
def test_block_lu_matches_lapack(P):
    print(f"case: block LU P={P}", flush=True)
    m = 3 * P
    rng = np.random.default_rng(P)
    A = rng.standard_normal((4, m, m))
    KB, ipiv, perm, info = block_lu(A)
    lu, piv, inf = KB.cpu().numpy(), ipiv.cpu().numpy(), info.cpu().numpy()
    assert np.all(inf == 0)
    for b in range(4):
        lu_ref, piv_ref = sl.lu_factor(A[b])
        assert np.array_equal(piv[b], piv_ref)
        assert rel_err(lu[b], lu_ref) < 1e-12
        # perm: row i of the factors is row perm[i] of the block
        Lm = np.tril(lu[b], -1) + np.eye(m)
        Um = np.triu(lu[b])
        assert rel_err(Lm @ Um, A[b][perm.cpu().numpy()[b]]) < 1e-13


def test_block_lu_reports_a_zero_pivot():
    A = np.stack([np.eye(96), np.eye(96)])
    A[1, 40, 40] = 0.0
    _, _, _, info = block_lu(A)
    assert info.cpu().numpy().tolist() == [0, 41]


@pytest.mark.parametrize("P", [1, 32, 128, 170])
def test_block_apply_solves_every_block(P):
    import torch
    print(f"case: block apply P={P}", flush=True)
    m, nblk = 3 * P, 6
    N = nblk * P
    rng = np.random.default_rng(7 + P)
    A = rng.standard_normal((nblk, m, m)) + 4 * np.eye(m)
    KB, _, perm, _ = block_lu(A)
    v = rng.standard_normal((3, N))
    z = torch.empty((3, N), dtype=torch.float64, device="cuda")
    check(lib().fl_block_apply(N, P, KB.data_ptr(), perm.data_ptr(), dev(v).data_ptr(), z.data_ptr(), None),
          "fl_block_apply")
    zh = z.cpu().numpy()
    for b in range(nblk):
        rhs = v[:, b * P:(b + 1) * P].T.reshape(-1)          # [point, component] order of the block
        want = np.linalg.solve(A[b], rhs)
        got = zh[:, b * P:(b + 1) * P].T.reshape(-1)
        # backward stable (the LU's normwise backward error), and as close to
        # LAPACK's answer as the conditioning allows
        berr = np.linalg.norm(A[b] @ got - rhs) / (np.linalg.norm(A[b], 2) * np.linalg.norm(got))
        assert berr < 1e-15 * m
        assert rel_err(got, want) < 1e-15 * np.linalg.cond(A[b])


@pytest.mark.parametrize("name,kernel", [("wall", fl.WallStokeslet3D()), ("brinkman:3.0", fl.Brinkmanlet3D(3.0))])
def test_assemble_blocks_are_the_diagonal_blocks(name, kernel):
    import torch
    X = rod_carpet(n_side=2, per_rod=16)
    N, P = len(X), 16
    KB = torch.empty((N // P, 3 * P, 3 * P), dtype=torch.float64, device="cuda")
    e = dev(np.array([0.01]))
    check(lib().fl_assemble_blocks(kernel_struct(kernel), N, P, dev(X.T).data_ptr(), e.data_ptr(), 0, 2.0,
                                   KB.data_ptr(), None), "fl_assemble_blocks")
    K = O.assemble(name, X, X, 0.01, mu=2.0)
    got = KB.cpu().numpy()
    for b in range(N // P):
        s = slice(3 * b * P, 3 * (b + 1) * P)
        assert rel_err(got[b], K[s, s]) < 1e-13


@pytest.mark.parametrize("spacing", [0.2, 5.0])
def test_gmres_with_block_jacobi_matches_lu_in_fewer_iterations(spacing):
    """Rod carpets (32 points per rod) in background shear: plain and
    block-Jacobi GMRES agree with the oracle's dense solve, and the
    preconditioned run takes fewer iterations: a few for well-separated rods,
    whose coupling is weak, and fewer than plain GMRES for the tight carpet."""
    print(f"case: block-Jacobi GMRES rods spacing {spacing}", flush=True)
    X = rod_carpet(n_side=4 if spacing < 1 else 3, spacing=spacing, per_rod=32)
    U = np.zeros_like(X)
    U[:, 0] = -X[:, 2]
    K = O.assemble("wall", X, X, 0.02)
    ref = np.linalg.solve(K, U.reshape(-1)).reshape(-1, 3)
    f0, info0 = fl.resistance(fl.WallStokeslet3D(), X, U, 0.02, method="gmres", rtol=1e-10, restart=50)
    f1, info1 = fl.resistance(fl.WallStokeslet3D(), X, U, 0.02, method="gmres", rtol=1e-10, restart=50,
                              block_points=32)
    assert rel_err(f0, ref) < 1e-6 and rel_err(f1, ref) < 1e-6
    assert info1.residual < 1e-10
    assert info1.iterations < info0.iterations
    if spacing > 1:
        assert info1.iterations * 3 <= info0.iterations
    with pytest.raises(ValueError):
        fl.resistance(fl.WallStokeslet3D(), X, U, 0.02, method="gmres", block_points=33)
