"""Published numbers that the reference's examples reproduce, through the
device path:

- the Gallagher-Smith biflagellate swims 5.75e-2 flagellar lengths per beat
  (5.529e-2 with the published code's shorter flagellum) through ``swim``
  (scipy RK45, rtol 1e-6) over device swimming problems
  (examples/biflagellate_swimmer/README.md:15-17, run.py:32);
- the rotational resistance R33 of a sphere on the six-patch grid matches
  Cortez, Fauci & Medovikov (2005) Table II to the two decimals printed
  there (examples/translating_sphere/run.py:28-30, README.md:9-14);
- a sphere in a Brinkman medium: summed force and torque against the closed
  forms 6 pi a (1 + alpha a + alpha^2 a^2 / 3) and
  8 pi a^3 (1 + alpha a + alpha^2 a^2 / 3) / (1 + alpha a) with the paper's
  blob widths on 7,996 points, to the errors the example reports
  (examples/sphere_in_brinkman/run.py:34-35, 46-51, README.md:11-15).
"""
import numpy as np
import pytest

import fluidlets_amd as fl
from fluidlets_amd.geometry import cube_sphere, fibonacci_sphere, gs_biflagellate

pytestmark = pytest.mark.gpu


# The published figures are what the reference's procedure gives: RK45 at
# rtol 1e-6 over one beat.  At that tolerance the step choices alone move the
# distance per beat by a few 1e-5 (solves that differ in the 13th digit gave
# 5.5288e-2 and 5.5313e-2 for the authors' shape), so the test follows the
# same procedure and allows 5e-5.  Converged in time (rtol 1e-8) the two
# shapes swim 5.737e-2 and 5.528e-2 per beat.
@pytest.mark.parametrize("variant,published,tol", [("paper", 5.75e-2, 5e-5), ("authors_code", 5.529e-2, 5e-5)])
def test_biflagellate_swim_speed(variant, published, tol):
    print(f"case: biflagellate one beat, {variant}", flush=True)
    from fluidlets_amd.swimming import swim
    shape_at = gs_biflagellate(variant)
    tr = swim(shape_at, (0.0, 2 * np.pi), 0.0125, 1.0, t_eval=[0.0, 2 * np.pi], rtol=1e-6, atol=1e-10)
    d = tr.x0[-1] - tr.x0[0]
    assert abs(np.linalg.norm(d) - published) < tol, np.linalg.norm(d)
    assert d[1] > 0 and abs(d[0]) < 1e-10 and abs(d[2]) < 1e-10


# Table II of Cortez, Fauci & Medovikov (2005): R33 for (eps, n) on 6 n^2 points
CFM_R33 = {(0.1, 12): 27.09, (0.05, 12): 25.53, (0.01, 12): 19.62,
           (0.1, 24): 27.16, (0.05, 24): 26.08, (0.01, 24): 23.89}


@pytest.mark.parametrize("n", [12, 24])
def test_sphere_rotational_resistance_matches_cfm_table(n):
    print(f"case: CFM Table II, n = {n}", flush=True)
    X = cube_sphere(n)
    for eps in (0.1, 0.05, 0.01):
        R, _, _ = fl.grand_resistance(fl.Stokeslet3D(), X, (0.0, 0.0, 0.0), eps)
        assert abs(R[5, 5] - CFM_R33[(eps, n)]) <= 0.005 + 1e-9, (eps, R[5, 5])
        # the sphere's symmetry: isotropic rotation, no coupling to translation
        assert np.ptp(np.diag(R)[3:]) < 1e-6 * R[5, 5] and np.abs(R[:3, 3:]).max() < 1e-6 * R[5, 5]


PAPER_EPS_DRAG = {0.1: 0.016, 1: 0.018, 5: 0.029, 10: 0.0373, 20: 0.0454}
PAPER_EPS_TORQUE = {0.1: 0.014, 1: 0.0055, 5: 0.0024, 10: 0.0030, 20: 0.00425}


@pytest.mark.parametrize("alpha", [0.1, 1.0, 5.0, 10.0, 20.0])
def test_brinkman_sphere_force_and_torque(alpha):
    print(f"case: Brinkman sphere alpha = {alpha}", flush=True)
    X = fibonacci_sphere(7996)
    K = fl.Brinkmanlet3D(alpha)
    c = (0.0, 0.0, 0.0)
    drag = 6 * np.pi * (1 + alpha + alpha**2 / 3)
    torque = 8 * np.pi * (1 + alpha + alpha**2 / 3) / (1 + alpha)
    key = alpha if alpha != int(alpha) else int(alpha)
    R, _, _ = fl.grand_resistance(K, X, c, PAPER_EPS_DRAG[key], precision="mixed", tol=1e-12)
    Rt, _, _ = fl.grand_resistance(K, X, c, PAPER_EPS_TORQUE[key], precision="mixed", tol=1e-12)
    # the example's table: force errors +0.1% to +3.6%, torque errors +0.3% to +1.0%
    assert 0.0 < R[2, 2] / drag - 1 < 0.037, R[2, 2] / drag - 1
    assert 0.0 < Rt[5, 5] / torque - 1 < 0.011, Rt[5, 5] / torque - 1
