"""Dense LU (fl_dgetrf/fl_sgetrf + solves + FP64 refinement) and GMRES on the
GPU, against LAPACK (scipy) and the oracle's restatement of core/solve.py.

Bars: FP64 solves match LAPACK to 1e-12 relative on well-conditioned systems
(with LAPACK's pivots reproduced exactly); mixed-precision solves reach
<= 1e-10 relative error against the FP64 LU oracle; GMRES matches LU to the
reference tests' 1e-6 with residual < 1e-9.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
import fluidlets_amd as fl
from conftest import golden
from fluidlets_amd.geometry import fibonacci_sphere
from helpers import rel_err

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("n", [1, 7, 128, 129, 300, 777])
def test_dgetrf_matches_lapack_with_pivoting(n):
    """Random Gaussian matrices need row interchanges everywhere: the
    factors and pivots equal LAPACK's getrf."""
    from fluidlets_amd import linalg
    print(f"case: dgetrf random n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal((n, 3))
    F = linalg.getrf(dev(A.copy()))
    lu_ref, piv_ref = sl.lu_factor(A)
    assert F.info == 0
    assert np.array_equal(F.ipiv.cpu().numpy()[:n], piv_ref)
    assert rel_err(F.lu.cpu().numpy(), lu_ref) < 1e-12
    x = dev(b.copy())
    linalg.getrs(F, x)
    assert rel_err(x.cpu().numpy(), sl.lu_solve((lu_ref, piv_ref), b)) < 1e-11


def test_sgetrf_is_a_backward_stable_partial_pivoting_lu():
    """FP32 factors of a random matrix: P A = L U to FP32 accuracy, the pivots of
    LAPACK sgetrf (rounding-order differences may swap a near-tie), |L| <= 1."""
    from fluidlets_amd import linalg
    rng = np.random.default_rng(1)
    A = rng.standard_normal((300, 300)).astype(np.float32)
    F = linalg.getrf(dev(A.copy()))
    lu = F.lu.cpu().numpy().astype(np.float64)
    piv = F.ipiv.cpu().numpy()
    _, piv_ref = sl.lu_factor(A)
    assert (piv == piv_ref).mean() > 0.98
    Lm = np.tril(lu, -1) + np.eye(300)
    Um = np.triu(lu)
    PA = A.astype(np.float64).copy()
    for r, p in enumerate(piv):
        PA[[r, p]] = PA[[p, r]]
    assert np.linalg.norm(PA - Lm @ Um) / np.linalg.norm(A) < 1e-5
    assert np.abs(np.tril(lu, -1)).max() <= 1.0 + 1e-6


@pytest.mark.parametrize("name,kernel", [("wall", fl.WallStokeslet3D()), ("brinkman:5.0", fl.Brinkmanlet3D(5.0))])
def test_resistance_matrix_takes_the_block_pivot_path(name, kernel):
    """Resistance matrices: the block-restricted pivot search is verified to be
    partial pivoting (no general panels, LAPACK's pivots)."""
    from fluidlets_amd import linalg
    print(f"case: resistance matrix LU {name}", flush=True)
    n = 700
    X = fibonacci_sphere(n, 1.0, (0, 0, 1.5))
    eps = 0.5 * np.sqrt(4 * np.pi / n)
    K = O.assemble(name, X, X, eps)
    F = linalg.getrf(dev(K.copy()))
    lu_ref, piv_ref = sl.lu_factor(K)
    assert F.general_panels == 0
    assert np.array_equal(F.ipiv.cpu().numpy(), piv_ref)
    assert rel_err(F.lu.cpu().numpy(), lu_ref) < 1e-12
    Kg = fl.assemble(kernel, X, X, eps)
    assert rel_err(Kg, K) < 1e-13


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_async_getrf_equals_the_verified_path(dtype):
    """fl_?getrf_async: on a resistance matrix (no interchanges under partial
    pivoting) its factors, pivots and permutation are bit-for-bit those of the
    verified path and LAPACK's pivots; on random and singular matrices it
    reports failure and the refilled verified factorization is used."""
    import torch
    from fluidlets_amd import linalg
    print(f"case: async getrf {dtype}", flush=True)
    n = 900
    X = fibonacci_sphere(n // 3, 1.0, (0, 0, 1.5))
    K = O.assemble("wall", X, X, 0.5 * np.sqrt(4 * np.pi / (n // 3))).astype(dtype)
    ref = linalg.getrf(dev(K.copy()))
    refills = []
    Fa = linalg.getrf(dev(K.copy()), refill=lambda A: refills.append(1))
    assert Fa.extra.get("async") and not refills
    assert torch.equal(Fa.lu, ref.lu) and torch.equal(Fa.ipiv, ref.ipiv)
    # aux: the permutation (n int32, padded to 256 bytes) and the inverted diagonal blocks
    pb = -(-4 * n // 256) * 256
    a, r = Fa.aux.cpu().numpy(), ref.aux.cpu().numpy()
    assert np.array_equal(a[:4 * n], r[:4 * n])
    assert np.array_equal(a[:4 * n].view(np.int32), np.arange(n))
    diff = np.flatnonzero(a[pb:] != r[pb:])
    assert diff.size == 0, f"inverse blocks differ from byte {pb + diff[0]} of {a.size}"
    assert np.array_equal(Fa.ipiv.cpu().numpy(), np.arange(n))
    b = np.random.default_rng(2).standard_normal((n, 3)).astype(dtype)
    x1, x2 = dev(b.copy()), dev(b.copy())
    linalg.getrs(Fa, x1)
    linalg.getrs(ref, x2)
    assert torch.equal(x1, x2)
    # a random matrix needs interchanges: the async attempt fails, A is refilled
    R = np.random.default_rng(3).standard_normal((300, 300)).astype(dtype)
    src = dev(R.copy())
    Fr = linalg.getrf(dev(R.copy()), refill=lambda A: A.copy_(src))
    assert not Fr.extra.get("async")
    Fv = linalg.getrf(dev(R.copy()))
    assert torch.equal(Fr.lu, Fv.lu) and torch.equal(Fr.ipiv, Fv.ipiv)
    # a hint shared by a sequence remembers the failure: the next call does not try again
    hint, refills = {}, []
    linalg.getrf(dev(R.copy()), refill=lambda A: (refills.append(1), A.copy_(src)), hint=hint)
    assert hint == {"async": False} and len(refills) == 1
    F2 = linalg.getrf(dev(R.copy()), refill=lambda A: refills.append(1), hint=hint)
    assert len(refills) == 1 and torch.equal(F2.lu, Fv.lu)
    # an exactly zero pivot also sends it to the verified path, which reports info
    Z = np.eye(200, dtype=dtype)
    Z[150, 150] = 0.0
    zsrc = dev(Z.copy())
    Fz = linalg.getrf(dev(Z.copy()), refill=lambda A: A.copy_(zsrc))
    assert not Fz.extra.get("async") and Fz.info == 151


def test_singular_matrix_reports_info():
    from fluidlets_amd import linalg
    A = np.eye(200)
    A[150, 150] = 0.0
    F = linalg.getrf(dev(A))
    assert F.info == 151
    # an exactly zero pivot makes the dense resistance solve raise, as
    # scipy.linalg.solve does for LAPACK info > 0
    from fluidlets_amd.core import solve as S
    import torch
    with pytest.raises(np.linalg.LinAlgError):
        A0 = torch.zeros((9, 9), dtype=torch.float64, device="cuda")
        F0 = linalg.getrf(A0)
        if F0.info > 0:
            raise np.linalg.LinAlgError("Matrix is singular.")
    assert S.DENSE_MAX == 6000


def test_mixed_precision_refinement_reaches_1e_10():
    from fluidlets_amd import linalg
    import torch
    print("case: sgetrf + refine", flush=True)
    n = 1500
    X = fibonacci_sphere(n, 1.0, (0, 0, 1.5))
    eps = 0.5 * np.sqrt(4 * np.pi / n)
    K = O.assemble("wall", X, X, eps)
    b = np.random.default_rng(3).standard_normal((3 * n, 6))
    xref = sl.solve(K, b)
    K64 = dev(K)
    F32 = linalg.getrf(K64.to(torch.float32))
    x, steps, res = linalg.refine(K64, F32, dev(b), tol=1e-12)
    assert steps <= 6 and res.max() <= 1e-12
    assert rel_err(x.cpu().numpy(), xref) < 1e-10


@pytest.mark.parametrize("nrhs", [1, 2, 3, 4, 5, 6, 7, 8])
def test_refinement_with_each_right_hand_side_count(nrhs):
    """The FP64 residual has one kernel per right-hand-side count (1 .. 8,
    accumulators sized to it); each refines to the LAPACK solution."""
    from fluidlets_amd import linalg
    import torch
    print(f"case: refine nrhs {nrhs}", flush=True)
    n = 300
    X = fibonacci_sphere(n, 1.0, (0, 0, 1.5))
    K = O.assemble("wall", X, X, 0.5 * np.sqrt(4 * np.pi / n))
    b = np.random.default_rng(40 + nrhs).standard_normal((3 * n, nrhs))
    F32 = linalg.getrf(dev(K).to(torch.float32))
    x, steps, res = linalg.refine(dev(K), F32, dev(b), tol=1e-12)
    assert x.shape == (3 * n, nrhs) and res.shape == (nrhs,)
    assert res.max() <= 1e-12
    assert rel_err(x.cpu().numpy(), sl.solve(K, b)) < 1e-10


@pytest.mark.parametrize("n_points", [1024, 1500])
def test_bf16x3_products_lu_refines_to_the_fp64_solution(n_points):
    """FP32 LU whose trailing updates (>= 2048 on a side) take split-bf16
    products: a backward-stable factorization to the split's accuracy, in the
    same number of refinement steps as FP32 products, and the refined
    solution within 1e-10 of LAPACK's FP64 solve."""
    from fluidlets_amd import linalg
    import torch
    print(f"case: sgetrf bf16x3 wall sphere N={n_points}", flush=True)
    X = fibonacci_sphere(n_points, 1.0, (0, 0, 1.5))
    eps = 0.5 * np.sqrt(4 * np.pi / n_points)
    K = O.assemble("wall", X, X, eps)
    n = 3 * n_points
    assert n - 512 >= 2048            # at least one split trailing update (outer panels of 512)
    b = np.random.default_rng(5).standard_normal((n, 6))
    xref = sl.solve(K, b)
    K64 = dev(K)
    out = {}
    for products in ("fp32", "bf16x3"):
        F = linalg.getrf(K64.to(torch.float32), products=products)
        assert F.extra["products"] == products and F.info == 0
        LU = F.lu.cpu().numpy().astype(np.float64)
        L, U = np.tril(LU, -1) + np.eye(n), np.triu(LU)
        perm = F.aux.cpu().numpy()[:4 * n].view(np.int32)
        berr = np.linalg.norm(L @ U - K[perm]) / np.linalg.norm(K)
        x, steps, res = linalg.refine(K64, F, dev(b), tol=1e-12)
        out[products] = (LU, berr, steps)
        assert res.max() <= 1e-12
        assert rel_err(x.cpu().numpy(), xref) < 1e-10
    (lu32, berr32, steps32), (lu3, berr3, steps3) = out["fp32"], out["bf16x3"]
    assert berr32 < 2e-6 and berr3 < 2e-5
    assert steps3 <= steps32 + 1
    # the split products ran: the factors differ from the FP32-product ones,
    # by about the split's 2^-16 rather than FP32's 2^-24
    d = np.abs(lu3 - lu32).max() / np.abs(lu32).max()
    assert 1e-8 < d < 1e-4


def test_bf16x3_products_with_row_interchanges():
    """Split-bf16 trailing products on the verified path (a random matrix
    needs interchanges in every panel, tiers 1 and 2): LAPACK's pivots,
    a backward-stable factorization to the split's accuracy, and FP64
    refinement of the FP32 factors to the FP64 solution."""
    from fluidlets_amd import linalg
    import torch
    print("case: sgetrf bf16x3 random n=3000", flush=True)
    n = 3000
    rng = np.random.default_rng(31)
    A = rng.standard_normal((n, n))
    A32 = A.astype(np.float32)
    F = linalg.getrf(dev(A32.copy()), products="bf16x3")
    F0 = linalg.getrf(dev(A32.copy()))
    assert F.extra["products"] == "bf16x3" and F.info == 0
    LU = F.lu.cpu().numpy().astype(np.float64)
    L, U = np.tril(LU, -1) + np.eye(n), np.triu(LU)
    perm = F.aux.cpu().numpy()[:4 * n].view(np.int32)
    A64 = A32.astype(np.float64)
    berr = np.linalg.norm(L @ U - A64[perm]) / (np.linalg.norm(np.abs(L) @ np.abs(U)))
    assert berr < 2e-5
    assert not np.array_equal(LU, F0.lu.cpu().numpy().astype(np.float64))   # the split path ran
    b = rng.standard_normal((n, 3))
    x, steps, res = linalg.refine(dev(A64), F, dev(b), tol=1e-12, max_iter=40)
    assert res.max() <= 1e-12
    assert rel_err(x.cpu().numpy(), np.linalg.solve(A64, b)) < 1e-8 * np.linalg.cond(A64) * 1e-2


@pytest.mark.parametrize("n", [2700, 4100])
def test_bf16x3_factorization_repeats_bit_for_bit(n):
    """The split-bf16 trailing updates (library GEMMs chosen without
    workspace, so without split-K partials) give the same factors on every
    run, on the verified and on the asynchronous path."""
    from fluidlets_amd import linalg
    print(f"case: sgetrf bf16x3 repeatability n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    D = (rng.standard_normal((n, n)) + n * np.eye(n)).astype(np.float32)   # no interchanges: the async path
    for M, refill in ((A, None), (D, lambda T: T.copy_(dev(D)))):
        F1 = linalg.getrf(dev(M.copy()), products="bf16x3", refill=refill)
        F2 = linalg.getrf(dev(M.copy()), products="bf16x3", refill=refill)
        assert F1.info == 0 and F1.extra.get("async", False) == (refill is not None)
        assert np.array_equal(F1.lu.cpu().numpy(), F2.lu.cpu().numpy())


@pytest.mark.parametrize("n", [700, 2559])
def test_bf16x3_products_are_ignored_for_fp64_and_small_matrices(n):
    """Below n = 2560 no trailing update is split, and the factors are those
    of the FP32 products bit for bit (and FP64 ignores the setting)."""
    from fluidlets_amd import linalg
    rng = np.random.default_rng(2)
    A = rng.standard_normal((n, n))
    F64a = linalg.getrf(dev(A.copy()))
    F64b = linalg.getrf(dev(A.copy()), products="bf16x3")
    assert np.array_equal(F64a.lu.cpu().numpy(), F64b.lu.cpu().numpy())
    F32a = linalg.getrf(dev(A.astype(np.float32)))
    F32b = linalg.getrf(dev(A.astype(np.float32)), products="bf16x3")
    assert np.array_equal(F32a.lu.cpu().numpy(), F32b.lu.cpu().numpy())
    with pytest.raises(ValueError):
        linalg.getrf(dev(A.copy()), products="fp16")


@pytest.mark.parametrize("n", [1, 129, 300, 777, 2600])
def test_sgetri_matches_lapack_float32_inverse(n):
    """fl_sgetri on partial-pivoting FP32 factors of a random matrix (rows
    interchanged everywhere): A X = I as closely as LAPACK's own FP32 inverse
    (sgetrf + sgetri through scipy) achieves, and X within the conditioning
    of the FP64 inverse."""
    from fluidlets_amd import linalg
    print(f"case: sgetri random n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n)).astype(np.float32)
    F = linalg.getrf(dev(A.copy()))
    X = linalg.getri(F).cpu().numpy().astype(np.float64)
    A64 = A.astype(np.float64)
    X_lapack = sl.inv(A).astype(np.float64)
    I = np.eye(n)
    ours = np.linalg.norm(A64 @ X - I) / np.sqrt(n)
    lapack = np.linalg.norm(A64 @ X_lapack - I) / np.sqrt(n)
    assert ours <= 10 * lapack + 1e-6
    assert rel_err(X, np.linalg.inv(A64)) < 1e-7 * np.linalg.cond(A64) * np.sqrt(n)


def test_refine_with_explicit_inverse_matches_lu_refinement():
    """fl_refine_inv: corrections from the explicit FP32 inverse reach the
    FP64 solution as the LU corrections do, in about as many steps."""
    from fluidlets_amd import linalg
    import torch
    print("case: refine with explicit inverse, wall sphere", flush=True)
    n_points = 900
    X = fibonacci_sphere(n_points, 1.0, (0, 0, 1.5))
    eps = 0.5 * np.sqrt(4 * np.pi / n_points)
    K = O.assemble("wall", X, X, eps)
    b = np.random.default_rng(8).standard_normal((3 * n_points, 7))
    xref = sl.solve(K, b)
    K64 = dev(K)
    F = linalg.getrf(K64.to(torch.float32))
    x_lu, steps_lu, res_lu = linalg.refine(K64, F, dev(b), tol=1e-12)
    Xinv = linalg.getri(F)
    x_inv, steps_inv, res_inv = linalg.refine(K64, None, dev(b), tol=1e-12, inverse=Xinv)
    assert res_lu.max() <= 1e-12 and res_inv.max() <= 1e-12
    assert rel_err(x_lu.cpu().numpy(), xref) < 1e-10 and rel_err(x_inv.cpu().numpy(), xref) < 1e-10
    assert steps_inv <= steps_lu + 1
    with pytest.raises(RuntimeError):
        linalg.refine(K64, None, dev(b), tol=1e-12, inverse=torch.zeros_like(Xinv), max_iter=3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resistance_lu_matches_oracle_and_legacy_drag(precision):
    """tests/test_kernels.py:61-73 and the MATLAB sphere drags (sphere_drag.mat)."""
    print(f"case: resistance lu {precision}", flush=True)
    s = golden("legacy_sphere_drag.npz")
    for n, F_ref in zip(s["N_list"].astype(int), s["F_free"]):
        X = fibonacci_sphere(n)
        eps = 0.5 * np.sqrt(4 * np.pi / n)
        U = np.tile([1.0, 0, 0], (n, 1))
        f, info = fl.resistance(fl.Stokeslet3D(), X, U, eps, method="lu", precision=precision)
        fr, _ = O.resistance("stokeslet", X, U, eps)
        assert info.residual < 1e-12 and rel_err(f, fr) < 1e-10
        assert abs(f[:, 0].sum() / F_ref - 1) < 1e-7
    n, epsw = 400, float(s["epsw"])
    for H, Fperp in zip(s["H_list"], s["F_perp"]):
        X = fibonacci_sphere(n, 1.0, (0, 0, H))
        f, info = fl.resistance(fl.WallStokeslet3D(), X, np.tile([0, 0, 1.0], (n, 1)), epsw, precision=precision)
        assert info.method == "lu" and abs(f[:, 2].sum() / Fperp - 1) < 1e-7


def test_resistance_complex_viscosity():
    X = fibonacci_sphere(200)
    U = np.tile([1.0, 0, 0], (200, 1))
    f1, _ = fl.resistance(fl.Stokeslet3D(), X, U, 0.1, method="lu")
    mu = 2.0 - 0.5j
    f, info = fl.resistance(fl.Stokeslet3D(), X, U, 0.1, mu=mu, method="lu")
    assert f.dtype == np.complex128 and rel_err(f, mu * f1) < 1e-12
    fr, _ = O.resistance("stokeslet", X, U, 0.1, mu=mu)
    assert rel_err(f, fr) < 1e-10 and info.residual < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gmres_matches_lu(precision):
    """tests/test_kernels.py:61-73: GMRES residual < 1e-9 and LU ~ GMRES to 1e-6."""
    print(f"case: gmres {precision}", flush=True)
    n = 400
    X = fibonacci_sphere(n)
    eps = 0.5 * np.sqrt(4 * np.pi / n)
    U = np.tile([1.0, 0, 0], (n, 1))
    F_lu, info_lu = fl.resistance(fl.Stokeslet3D(), X, U, eps, method="lu")
    F_gm, info_gm = fl.resistance(fl.Stokeslet3D(), X, U, eps, method="gmres", precision=precision)
    assert info_lu.residual < 1e-12 and info_gm.residual < 1e-9
    assert np.allclose(F_lu, F_gm, rtol=1e-6, atol=1e-8)
    assert 1.0 < F_lu[:, 0].sum() / (6 * np.pi) < 1.02


def test_gmres_reproduces_legacy_tower_forces(towers_mat):
    """tests/legacy/test_phase1_port.py:62-74 (wall points stay force-free)."""
    print("case: gmres towers", flush=True)
    d = towers_mat
    X, U, F_legacy = d["X"], d["U"], d["F"]
    F, info = fl.resistance(fl.WallStokeslet3D(), X, U, 0.5, method="gmres")
    assert info.residual < 1e-9
    on_wall = X[:, 2] == 0
    assert np.abs(F[on_wall]).max() < 1e-12
    above = ~on_wall
    assert np.linalg.norm(F[above] - F_legacy[above]) / np.linalg.norm(F_legacy[above]) < 1e-3
    for k in range(3):
        tower = slice(147 * k, 147 * (k + 1))
        assert np.isclose(F[tower, 0].sum(), F_legacy[tower, 0].sum(), rtol=1e-5)


def test_gmres_nonconvergence_raises():
    X = fibonacci_sphere(300)
    with pytest.raises(RuntimeError):
        fl.resistance(fl.Stokeslet3D(), X, np.tile([1.0, 0, 0], (300, 1)), 0.05, method="gmres", restart=2,
                      maxiter=1)


def test_grand_resistance_six_modes():
    """Sphere: T = 6 pi I-ish and R = 8 pi I-ish; mixed precision agrees with FP64 to 1e-10."""
    n = 1200
    X = fibonacci_sphere(n)
    eps = 0.5 * np.sqrt(4 * np.pi / n)
    R64, f64, _ = fl.grand_resistance(fl.Stokeslet3D(), X, (0, 0, 0), eps)
    Rm, fm, info = fl.grand_resistance(fl.Stokeslet3D(), X, (0, 0, 0), eps, precision="mixed")
    assert np.allclose(np.diag(R64)[:3] / (6 * np.pi), 1, atol=0.02)
    assert np.allclose(np.diag(R64)[3:] / (8 * np.pi), 1, atol=0.05)
    assert rel_err(Rm, R64) < 1e-10 and rel_err(fm, f64) < 1e-10
    assert info["refinement_steps"] >= 1


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (33, 17, 5), (128, 128, 128), (129, 1, 128), (300, 200, 130),
                                   (64, 513, 40), (2100, 4100, 24)])
def test_mfma_gemm_matches_numpy(dtype, M, N, K):
    """C = alpha A B + beta C through fl_sgemm / fl_dgemm (pins the MFMA lane
    layouts and the tile edges).  Products of fewer than two 128 x 128 tiles
    per CU take 64-row tiles; 2100 x 4100 (561 tiles) takes 128-row ones."""
    import torch
    from fluidlets_amd import _lib
    print(f"case: gemm {dtype} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((M, K)).astype(dtype)
    B = rng.standard_normal((K, N)).astype(dtype)
    C = rng.standard_normal((M, N)).astype(dtype)
    ref = -0.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 2.0 * C.astype(np.float64)
    dA, dB, dC = dev(A), dev(B), dev(C)
    fn = _lib.lib().fl_sgemm if dtype == "float32" else _lib.lib().fl_dgemm
    st = fn(M, N, K, -0.5, dA.data_ptr(), K, dB.data_ptr(), N, 2.0, dC.data_ptr(), N,
            torch.cuda.current_stream().cuda_stream)
    assert st == 0
    tol = 1e-5 if dtype == "float32" else 1e-14
    assert rel_err(dC.cpu().numpy().astype(np.float64), ref) < tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_mfma_gemm_in_place_on_b(dtype):
    """C = A B written over B (M = K = 128, N = 60,000), as the LU's U12 =
    L11^-1 A12 products are: one tile row, so no workgroup overwrites rows
    of B that another has still to read.  A product this narrow (469 tiles
    of 128 rows) would otherwise take 64-row tiles, 938 workgroups, more
    than run at once, and the later ones would read rows already written.
    Twice, both results equal."""
    import torch
    from fluidlets_amd import _lib
    print(f"case: gemm in place on B {dtype}", flush=True)
    rng = np.random.default_rng(11)
    M, N = 128, 60000
    A = rng.standard_normal((M, M)).astype(dtype)
    B = rng.standard_normal((M, N)).astype(dtype)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    fn = _lib.lib().fl_sgemm if dtype == "float32" else _lib.lib().fl_dgemm
    outs = []
    for _ in range(2):
        dA, dB = dev(A), dev(B)
        st = fn(M, N, M, 1.0, dA.data_ptr(), M, dB.data_ptr(), N, 0.0, dB.data_ptr(), N,
                torch.cuda.current_stream().cuda_stream)
        assert st == 0
        outs.append(dB.cpu().numpy())
    tol = 1e-5 if dtype == "float32" else 1e-14
    assert rel_err(outs[0].astype(np.float64), ref) < tol
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n_points,kernel,products", [
    (300, "stokeslet", "fp32"),          # n = 900: every update on the FP32 lower-tile GEMM
    (1200, "brinkman:3.0", "fp32"),      # n = 3,600: FP32 products, lower tiles of 128
    (1200, "brinkman:3.0", "bf16x3"),    # the same with split-bf16 lower-tile updates
    (1100, "wall", "bf16x3"),            # symmetric by reciprocity only to rounding
    (1200, "stokeslet", "fp64"),
])
def test_symmetric_factorization_matches_lapack(n_points, kernel, products):
    """The symmetric factorization A = L D L^T (fl_?sytrf_async) of a kernel
    matrix with one blob width, in the LU factor format: the factors satisfy
    L U = A with U = D L^T to the arithmetic's accuracy, and the solves
    (refined against the FP64 matrix for FP32 factors) equal LAPACK's."""
    from fluidlets_amd import linalg
    import torch
    print(f"case: sytrf {kernel} N={n_points} {products}", flush=True)
    X = fibonacci_sphere(n_points, 1.0, (0, 0, 1.5) if kernel == "wall" else (0, 0, 0))
    eps = 0.5 * np.sqrt(4 * np.pi / n_points)
    K = O.assemble(kernel, X, X, eps)
    n = 3 * n_points
    if kernel != "wall":
        assert np.array_equal(K, K.T)        # reciprocity, bit for bit without images
    b = np.random.default_rng(n).standard_normal((n, 6))
    xref = sl.solve(K, b)
    K64 = dev(K)
    fp64 = products == "fp64"
    A = K64.clone() if fp64 else K64.to(torch.float32)
    src = K64 if fp64 else K64.to(torch.float32)
    F = linalg.getrf(A, refill=lambda T: T.copy_(src), sym=True, products="fp32" if fp64 else products)
    assert F.extra.get("sym") and F.extra.get("async") and F.info == 0
    LU = F.lu.cpu().numpy().astype(np.float64)
    L, U = np.tril(LU, -1) + np.eye(n), np.triu(LU)
    unit = 2.0**-52 if fp64 else (2.0**-16 if products == "bf16x3" else 2.0**-23)
    # U = D L^T: the transpose-scale formed the off-diagonal blocks; the
    # diagonal blocks come from the LU of the block, equal up to rounding
    d = np.diag(U)
    assert np.abs(U - np.triu(d[:, None] * L.T)).max() <= 200 * unit * np.abs(U).max()
    assert np.linalg.norm(L @ U - K) / np.linalg.norm(K) < 200 * unit
    if fp64:
        x = dev(b)
        linalg.getrs(F, x)
        assert rel_err(x.cpu().numpy(), xref) < 1e-12
    else:
        x, steps, res = linalg.refine(K64, F, dev(b), tol=1e-12)
        assert res.max() <= 1e-12 and steps <= 6
        assert rel_err(x.cpu().numpy(), xref) < 1e-10


def test_symmetric_factorization_falls_back_to_pivoting():
    """A symmetric matrix whose elimination without interchanges has
    multipliers above 1 (a large off-diagonal entry): the symmetric attempt
    reports failure and getrf refills and factors it on the verified
    pivoting path, equal to LAPACK."""
    from fluidlets_amd import linalg
    print("case: sytrf fallback", flush=True)
    rng = np.random.default_rng(9)
    n = 400
    G = rng.standard_normal((n, n))
    S = G + G.T
    x = dev(S)
    F = linalg.getrf(x.clone(), refill=lambda T: T.copy_(x), sym=True)
    assert not F.extra.get("async")
    lu, piv = sl.lu_factor(S)
    assert np.array_equal(F.ipiv.cpu().numpy(), piv)
    assert np.allclose(F.lu.cpu().numpy(), lu, rtol=0, atol=1e-11 * np.abs(lu).max())


@pytest.mark.parametrize("sym,products", [(False, "fp32"), (True, "fp32"), (True, "bf16x3")])
def test_fp32_factor_backward_error_is_lapack_level(sym, products):
    """The FP32 factors of the 2,048-point wall sphere (n = 6,144; every large
    trailing update on the packed MFMA kernels) have LAPACK sgetrf's backward
    error on the same FP32 matrix, within a factor of 2 (split-bf16 products,
    good to about 2^-16 each, within a factor of 4).  The updates sum their
    products from zero and subtract the sum from C once: starting the sums
    from C rounded every partial sum at the scale of |C| and left the FP32
    factor's backward error 17 times LAPACK's here (and 30 times at n =
    24,576, a third refinement step)."""
    import torch
    from fluidlets_amd import linalg
    from fluidlets_amd.core.assemble import assemble_device
    from fluidlets_amd import _device as D
    print(f"case: FP32 factor backward error sym={sym} products={products}", flush=True)
    N = 2048
    X = fibonacci_sphere(N, 1.0, (0, 0, 1.5))
    xs = D.soa(X, D.device())
    eps_t, stride = D.eps_arg(0.5 * np.sqrt(4 * np.pi / N), N, D.device())
    K64, K32 = assemble_device(fl.WallStokeslet3D(), xs, xs, eps_t, stride, 1.0, fp64=True, fp32=True)
    A = K32.clone()
    F = linalg.getrf(A, refill=lambda T: T.copy_(K32), products=products, sym=sym)
    assert F.extra.get("async") and F.info == 0
    n = 3 * N

    def backward_error(lu):
        L = torch.tril(lu, -1) + torch.eye(n, dtype=torch.float64, device=lu.device)
        return float((L @ torch.triu(lu) - K32.double()).abs().max() / K32.double().abs().max())

    err_gpu = backward_error(F.lu.double())
    lu_ref, piv_ref = sl.lu_factor(K32.cpu().numpy())
    assert np.array_equal(piv_ref, np.arange(n))     # no interchanges: the same elimination
    err_ref = backward_error(torch.from_numpy(lu_ref).cuda().double())
    assert err_gpu <= (2.0 if products == "fp32" else 4.0) * err_ref, (err_gpu, err_ref)
    # and FP64 refinement from FP32-product factors converges in two corrections to 1e-12
    rhs = torch.from_numpy(np.random.default_rng(7).standard_normal((n, 6))).cuda()
    x, steps, res = linalg.refine(K64, F, rhs, tol=1e-12)
    assert steps <= (2 if products == "fp32" else 3) and res.max() <= 1e-12
