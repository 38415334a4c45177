"""The barycentric Lagrange treecode on the GPU (fluidlets_amd.treecode,
fl_bltc_weights / fl_bltc_eval) against the direct kernel sums, which the
oracle pins (test_gpu_kernel_sums.py).

- theta = 0 takes every source directly: the direct sum in another order,
  to rounding, for every kernel, real and complex viscosity and strengths;
- the proxy strengths reproduce each cluster's moments: barycentric
  interpolation at n + 1 Chebyshev points is exact for polynomials of degree
  <= n in each coordinate, so sum_k F_k s_k^a = sum_j f_j y_j^a for every
  monomial x^a1 y^a2 z^a3 with a_i <= n (a property of the method, not of
  this code);
- the far-field error falls exponentially with the degree and is below the
  stated bound at the defaults (degree 8, theta 0.7).
"""
import numpy as np
import pytest

import fluidlets_amd as fl
from fluidlets_amd.treecode import Treecode, evaluate_treecode
from helpers import rel_err

pytestmark = pytest.mark.gpu

KERNELS = [fl.WallStokeslet3D(), fl.Stokeslet3D(), fl.Brinkmanlet3D(4.0)]


def slab(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform([-4, -4, 0.05], [4, 4, 4], (n, 3)), rng.standard_normal((n, 3))


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: type(k).__name__)
def test_theta_zero_is_the_direct_sum(kernel):
    print(f"case: treecode theta 0 {type(kernel).__name__}", flush=True)
    X, F = slab(12000, 1)
    Y = np.random.default_rng(5).uniform([-5, -5, 0.05], [5, 5, 5], (3000, 3))
    for T, mu, f in ((X, 1.0, F), (Y, 1.0 + 0.5j, F), (X, 2.0, F + 1j * F[::-1])):
        u = fl.evaluate(kernel, T, X, f, 0.02, mu)
        v = evaluate_treecode(kernel, T, X, f, 0.02, mu, theta=0.0, leaf_size=500)
        assert v.dtype == u.dtype and v.shape == u.shape
        assert rel_err(v, u) < 1e-13


@pytest.mark.parametrize("degree", [1, 3, 6, 8, 10, 12])
def test_proxy_strengths_reproduce_the_cluster_moments(degree):
    """Degrees 1 .. 12 take 64, 16, 5, 3, 2 and 1 groups of lanes per
    workgroup (summed pairwise at the end); the largest clusters with
    proxies (5,000 sources) are summed in two chunks of up to 4,096."""
    import torch
    print(f"case: treecode proxy moments degree {degree}", flush=True)
    X, F = slab(20000, 2)
    tc = Treecode(fl.Stokeslet3D(), X, X, 0.02, degree=degree, theta=0.7, leaf_size=400)
    assert tc.ncl > 0
    f = torch.from_numpy(np.ascontiguousarray(F.T)).cuda()
    tc.apply_device(f)
    torch.cuda.synchronize()
    px = tc._px.cpu().numpy()
    pf = tc._pf.cpu().numpy()
    cl = tc._cl.cpu().numpy().reshape(-1, 5)
    assert cl[:, 4].max() == 2                        # clusters of 5,000 sources in two chunks
    Xs, Fs = X[tc.st.order], F[tc.st.order]
    nq = (degree + 1) ** 3
    for b, e, base, _, _ in cl[:: max(1, len(cl) // (20 if degree <= 6 else 4))]:
        y, fy = Xs[b:e], Fs[b:e]
        s, fs = px[:, base:base + nq].T, pf[:, base:base + nq].T
        c = 0.5 * (y.min(0) + y.max(0))
        h = 0.5 * (y.max(0) - y.min(0))
        for a in np.ndindex(degree + 1, degree + 1, degree + 1):
            my = np.prod(((y - c) / h) ** np.array(a), axis=1)
            ms = np.prod(((s - c) / h) ** np.array(a), axis=1)
            ref = my @ fy
            assert np.abs(ms @ fs - ref).max() <= 1e-11 * (np.abs(my) @ np.abs(fy)).max()


def test_error_falls_with_the_degree():
    print("case: treecode convergence", flush=True)
    X, F = slab(60000, 3)
    K = fl.WallStokeslet3D()
    u = fl.evaluate(K, X, X, F, 0.02)
    errs = [rel_err(evaluate_treecode(K, X, X, F, 0.02, degree=d, theta=0.7, leaf_size=1000), u) for d in (2, 4, 6, 8)]
    print("errors", errs, flush=True)
    assert all(b < a / 5 for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 2e-6


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: type(k).__name__)
def test_defaults_meet_the_stated_error(kernel):
    """Degree 8, theta 0.7, leaves of 1,024: relative error below 1e-5
    against the direct sum, for targets apart from the sources too."""
    print(f"case: treecode defaults {type(kernel).__name__}", flush=True)
    X, F = slab(80000, 4)
    T = np.random.default_rng(6).uniform([-6, -6, 0.05], [6, 6, 6], (20000, 3))
    for tg in (X, T):
        u = fl.evaluate(kernel, tg, X, F, 0.02)
        v = evaluate_treecode(kernel, tg, X, F, 0.02)
        assert rel_err(v, u) < 1e-5
