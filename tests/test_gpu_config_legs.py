"""The workloads' remaining legs at their full sizes on the GPU:

- the resistance solve of the 8,192-point sphere (n = 24,576, six rigid-body
  right-hand sides) with free-space Brinkmanlet3D at alpha = 1, 5, 10: the
  mixed-precision solve within 1e-10 of the FP64 LU, the FP64 forces
  reproducing the prescribed velocities through the kernel sum to 1e-12, and
  the grand resistance of a sphere (translation-rotation decoupled, drag
  rising with alpha like 6 pi a (1 + alpha a + alpha^2 a^2 / 3));
- the rod carpet (131,072 points) with flexible GMRES (FP32C inner cycles,
  FP64 outer loop) to 1e-8, against the FP64 GMRES solve;
- a serving mix drawn from the full serving distribution (ellipsoids up to
  6,000 points, the 1,000 + 2 x 400-point biflagellate stepped by RK2 in
  sessions): each result is bit for bit the same alone and in the mix, the
  server's grand resistances equal the library's, and a full-size swimmer
  step equals the oracle's dense bordered solve;
- twelve swimmers (each with its own helper threads and streams) served
  beside resistance solves: results alone and in the mix equal bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import fluidlets_amd as fl
from fluidlets_amd.geometry import fibonacci_sphere, rod_carpet
from fluidlets_amd.serving import Request, SolveServer, serving_mix
from helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha", [1.0, 5.0, 10.0])
def test_full_size_brinkman_sphere_mixed_equals_fp64(alpha):
    import torch
    print(f"case: Brinkmanlet({alpha}) sphere N=8192 mixed vs fp64", flush=True)
    N = 8192
    X = fibonacci_sphere(N, 1.0, (0.0, 0.0, 0.0))
    eps = 0.5 * np.sqrt(4 * np.pi / N)
    K = fl.Brinkmanlet3D(alpha)
    c = (0.0, 0.0, 0.0)
    R64, f64, _ = fl.grand_resistance(K, X, c, eps, precision="fp64")
    torch.cuda.empty_cache()
    Rm, fm, info = fl.grand_resistance(K, X, c, eps, precision="mixed", tol=1e-12)
    torch.cuda.empty_cache()
    assert rel_err(fm, f64) < 1e-10
    assert np.abs(Rm - R64).max() < 1e-12 * np.abs(R64).max()
    # a sphere: decoupled translation and rotation, isotropic drag
    d = np.diag(R64)
    assert np.all(d > 0)
    assert np.abs(R64[:3, 3:]).max() < 1e-3 * d[0] and np.abs(R64[3:, :3]).max() < 1e-3 * d[0]
    assert np.ptp(d[:3]) < 1e-3 * d[0] and np.ptp(d[3:]) < 1e-3 * d[3]
    # drag of a sphere in a Brinkman medium, to the discretization's reach
    # (the layer 1/alpha is 2.5 point spacings at alpha = 10)
    exact = 6 * np.pi * (1 + alpha + alpha**2 / 3)
    assert 0.8 * exact < d[0] < 1.05 * exact
    # the FP64 residual of one mode at rounding level (matrix-free, by the kernel sum)
    U = np.zeros((N, 3))
    U[:, 0] = 1.0
    assert rel_err(fl.evaluate(K, X, X, f64[0], eps), U) < 1e-12


def test_full_size_rod_carpet_flexible_gmres():
    """C4's flexible variant: FP64 outer GMRES(50) with FP32C inner cycles
    reaches 1e-8, and its forces agree with the FP64 GMRES solve to what two
    solutions of a 1e-8 residual can (measured 5e-7 on this carpet)."""
    from fluidlets_amd import _device as D
    from fluidlets_amd.core.gmres import gmres_device
    import torch
    print("case: rod carpet 131072 points flexible GMRES", flush=True)
    X = rod_carpet(n_side=32, per_rod=128)
    N = len(X)
    assert N == 131072
    dev = D.device()
    x_soa = D.soa(X, dev)
    eps_t, stride = D.eps_arg(0.01, N, dev)
    U = np.zeros((N, 3))
    U[:, 0] = -X[:, 2]
    U_soa = D.soa(U, dev)
    f64, it64, r64 = gmres_device(fl.WallStokeslet3D(), x_soa, eps_t, stride, U_soa, rtol=1e-8, restart=50,
                                  maxiter=40)
    fm, itm, rm = gmres_device(fl.WallStokeslet3D(), x_soa, eps_t, stride, U_soa, rtol=1e-8, restart=50,
                               maxiter=40, precision="mixed")
    assert r64 <= 1e-8 and rm <= 1e-8
    assert itm < it64               # outer steps: the inner cycles do most of the work
    err = (torch.linalg.norm(fm - f64) / torch.linalg.norm(f64)).item()
    assert err < 2e-6, err
    # the flexible solution's own residual, recomputed by an FP64 kernel sum
    u = fl.evaluate(fl.WallStokeslet3D(), X, X, fm.cpu().numpy().T, 0.01)
    assert rel_err(u, U) < 2e-8


def _c5_subset():
    """Six resistance requests from the 2,000-request serving trace (the
    largest and smallest N, both kernels) and the first steps of two of its
    swimmer sessions, arriving within a few milliseconds."""
    reqs = serving_mix(2000, seed=5, rate=1000.0)
    res = [r for r in reqs if r.kind == "resistance"]
    by_n = sorted(res, key=lambda r: len(r.payload["points"]))
    pick = [by_n[0], by_n[-1], by_n[len(by_n) // 2]]
    taken = {r.rid for r in pick}
    pick += [r for r in res if type(r.payload["kernel"]).__name__ == "Brinkmanlet3D" and r.rid not in taken][:2]
    taken = {r.rid for r in pick}
    pick += [r for r in res if type(r.payload["kernel"]).__name__ == "WallStokeslet3D" and r.rid not in taken][:1]
    steps = [r for r in reqs if r.kind == "swim_step" and r.payload["session"] in ("swimmer0", "swimmer1")
             and r.payload["seq"] < 2]
    out = [Request(r.kind, r.payload, 0.001 * k, r.rid) for k, r in enumerate(pick + steps)]
    return out


def test_c5_shaped_mix_alone_equals_mix_and_library():
    import torch
    print("case: serving mix from the full distribution (N <= 6000, 1800-point swimmers)", flush=True)
    mix = _c5_subset()
    assert max(len(r.payload["points"]) for r in mix if r.kind == "resistance") > 5500
    srv = SolveServer(slots=4, precision="mixed")
    out = {r.rid: r for r in srv.serve(mix, realtime=True)}
    srv.close()
    assert all(r.error is None for r in out.values()), [r.error for r in out.values()]
    alone = SolveServer(slots=1, precision="mixed")
    for r in [r for r in mix if r.kind == "resistance"]:
        ref = alone.serve([Request(r.kind, r.payload, 0.0, r.rid)], realtime=False)[0]
        assert ref.error is None and np.array_equal(ref.value, out[r.rid].value), r.rid
        p = r.payload
        R, _, _ = fl.grand_resistance(p["kernel"], p["points"], p["center"], p["eps"])
        assert np.allclose(out[r.rid].value, R, rtol=0, atol=1e-9 * np.abs(R).max()), r.rid
        torch.cuda.empty_cache()
    for key in ("swimmer0", "swimmer1"):
        steps = [r for r in mix if r.kind == "swim_step" and r.payload["session"] == key]
        ref = alone.serve([Request(r.kind, r.payload, 0.0, r.rid) for r in steps], realtime=False)
        for a, r in zip(ref, steps):
            assert a.error is None and np.array_equal(a.value, out[r.rid].value), (key, r.rid)
    alone.close()


def test_c5_swimmer_step_matches_the_oracle_bordered_solve():
    """The first right-hand side of a served swimmer session (the body at
    t = 0 in its seeded frame): U, Omega and the forces of the device
    Schur-complement solve against the oracle's dense (3N + 6) solve."""
    from fluidlets_amd.swimming import swimming_problem
    print("case: full-size swimmer (n = 5,406) against the oracle", flush=True)
    step = next(r for r in serving_mix(60, seed=5, rate=1000.0) if r.kind == "swim_step")
    spec = step.payload["spec"]
    shape = spec["shape_at"](0.0)
    assert len(shape.x) == 1800
    nearest = np.arange(len(shape.x))
    for precision in ("fp64", "mixed"):
        st = swimming_problem(shape, spec["x0"], spec["B0"], spec["eps"], nearest=nearest, precision=precision)
        U, W, f, _ = O.swimming_problem("stokeslet", shape.x, shape.v, shape.X, spec["x0"], spec["B0"], spec["eps"],
                                        nearest=nearest)
        scale = max(np.abs(U).max(), np.abs(W).max())
        assert np.abs(st.U - U).max() < 1e-10 * scale and np.abs(st.Omega - W).max() < 1e-10 * scale, precision
        assert rel_err(st.forces, f) < 1e-10, precision


def test_many_swimmers_with_helpers_alone_equal_mix():
    """Twelve swimmers, each with its own factor cache and three helper
    threads on streams of their own, beside resistance solves on four
    slots: every result equals the same request served alone, bit for bit
    (no two live users share a stream or a workspace)."""
    print("case: 12 swimmers with helpers beside resistance solves", flush=True)
    reqs = serving_mix(40, seed=11, n_min=200, n_max=1200, swimmer_frac=0.0, rate=1000.0)[:4]
    from fluidlets_amd.swimming import biflagellate
    shape_at = biflagellate(n_body=150, n_flag=40)
    for k in range(12):
        pay = {"shape_at": shape_at, "dt": 1.0 / 8, "steps": 3 + k % 3, "eps": 0.02, "n_points": 230}
        reqs.append(Request("swimmer", pay, 0.0, 100 + k))
    reqs = [Request(r.kind, r.payload, 0.001 * i, r.rid) for i, r in enumerate(reqs)]
    srv = SolveServer(slots=4, precision="mixed")
    out = {r.rid: r for r in srv.serve(reqs, realtime=True)}
    srv.close()
    assert all(r.error is None for r in out.values()), [r.error for r in out.values()]
    alone = SolveServer(slots=1, precision="mixed")
    for r in reqs:
        ref = alone.serve([Request(r.kind, r.payload, 0.0, r.rid)], realtime=False)[0]
        assert ref.error is None and np.array_equal(ref.value, out[r.rid].value), r.rid
    alone.close()
