"""The plain FP64 translating-sphere problem, the reference's own CPU case: a
level-3 icosphere (642 points) of radius 1, eps = 0.25 x the mean
nearest-neighbour spacing, Stokeslet3D, mu = 1, rigid velocity U = (1, 0, 0);
assemble the 1926 x 1926 matrix, solve by dense LU, total drag against
6 pi mu a U; then evaluate the solved forces at 10,000 targets uniform in the
shell 1.5 < r < 5 (seed 1).

Bars: forces and velocities within 1e-12 of the oracle's LAPACK solve and C
kernel sums; the physics (drag and the far field of the exact Stokes
solution) within the discretization's 3 %, which the oracle shows equally
(drag / 6 pi = 0.978, far-field error 2.2 %)."""
import numpy as np
import pytest

import oracle as O
import fluidlets_amd as fl
from fluidlets_amd.geometry import icosphere, mean_nearest_spacing
from helpers import rel_err

pytestmark = pytest.mark.gpu


def c1_problem():
    X = icosphere(3)
    eps = 0.25 * mean_nearest_spacing(X)
    U = np.tile([1.0, 0.0, 0.0], (len(X), 1))
    rng = np.random.default_rng(1)
    d = rng.normal(size=(10000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(1.5 ** 3, 5.0 ** 3, 10000) ** (1.0 / 3.0)
    return X, eps, U, d * r[:, None]


def stokes_sphere_flow(T, U, a=1.0):
    """Exact Stokes flow past a translating sphere of radius a (mu = 1)."""
    r = np.linalg.norm(T, axis=1)
    e = T / r[:, None]
    eu = (e @ U)[:, None] * e
    return (3 * a / (4 * r))[:, None] * (U + eu) + (a ** 3 / (4 * r ** 3))[:, None] * (U - 3 * eu)


def test_c1_translating_sphere():
    print("case: C1 icosphere drag and shell velocities", flush=True)
    X, eps, U, T = c1_problem()
    assert len(X) == 642 and abs(eps - 0.0352975) < 1e-6
    f, info = fl.resistance(fl.Stokeslet3D(), X, U, eps, method="lu")
    fr, _ = O.resistance("stokeslet", X, U, eps)
    assert info.method == "lu" and info.residual < 1e-12
    assert rel_err(f, fr) < 1e-12
    drag = f[:, 0].sum() / (6 * np.pi)
    assert abs(drag - 0.9784) < 1e-3 and abs(drag - 1) < 0.03
    u = fl.evaluate(fl.Stokeslet3D(), T, X, f, eps)
    ur = O.evaluate("stokeslet", T, X, fr, eps)
    assert rel_err(u, ur) < 1e-12
    assert rel_err(u, stokes_sphere_flow(T, U[0])) < 0.03
