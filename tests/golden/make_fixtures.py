"""Write the golden fixtures under tests/golden/ (run once, in a container that
has the reference checkout; the fixtures are committed, the script is kept
beside them so they can be regenerated).

Sources, all DATA files that the reference's own tests read (no reference code
is imported or run):

* tests/legacy/data/kernels.mat       -> legacy_kernels.npz
  (MATLAB StokesletEval.m / ImageEval.m outputs; asserted at
  tests/legacy/test_phase1_port.py:33-44 and test_phase0.py:28-46)
* tests/legacy/data/towers.mat        -> legacy_towers.npz
  (x0, Vel, Forces, resid of the legacy GMRES tower run;
  tests/legacy/test_phase1_port.py:47-59)
* legacy/sdc_cuda3/velocity_t_xyz_final.raw -> legacy_grid.npz
  (32^3 x 3 float32 velocity grid of the 2010 CUDA run;
  tests/legacy/test_phase1_port.py:77-87)
* tests/legacy/data/sphere_drag.mat   -> legacy_sphere_drag.npz
  (tests/legacy/test_phase0.py:87-105, tests/legacy/matlab/check_sphere_drag.m)

plus values this script computes itself with mpmath at 50 digits:

* brinkman_mp.npz: f1(z) = e^{-z}(1 + 1/z + 1/z^2) - 1/z^2 and
  f2(z) = 3/z^2 - e^{-z}(1 + 3/z + 3/z^2) (kernels/brinkmanlet3d.py:19-28),
  at the z of tests/test_brinkmanlet.py:32 and a denser sweep.

The data are BSD-3 licensed (LICENSE-fluidlets.txt in this folder).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")


def loadmat(name):
    import scipy.io as sio
    return sio.loadmat(REF / "tests" / "legacy" / "data" / name, squeeze_me=True)


def main():
    d = loadmat("kernels.mat")
    keys = ["targets", "sources", "strengths", "deltas", "V_stokeslet", "V_image", "wall",
            "V_wall_free", "V_wall_total", "wall_err", "eps_list"]
    np.savez_compressed(HERE / "legacy_kernels.npz", **{k: np.asarray(d[k], dtype=np.float64) for k in keys})

    t = loadmat("towers.mat")
    np.savez_compressed(HERE / "legacy_towers.npz",
                        x0=np.asarray(t["x0"], np.float64), Vel=np.asarray(t["Vel"], np.float64),
                        Forces=np.asarray(t["Forces"], np.float64), resid=np.float64(t["resid"]))

    raw = np.fromfile(REF / "legacy" / "sdc_cuda3" / "velocity_t_xyz_final.raw", "<f4")
    np.savez_compressed(HERE / "legacy_grid.npz", velocity=raw)

    s = loadmat("sphere_drag.mat")
    np.savez_compressed(HERE / "legacy_sphere_drag.npz",
                        **{k: np.asarray(s[k], dtype=np.float64) for k in
                           ["N_list", "h_list", "F_free", "F_stokes", "H_list", "F_par", "F_perp",
                            "F_free_N", "epsw"]})

    import mpmath
    mpmath.mp.dps = 50
    zs = [1e-8, 0.01, 0.3, 0.499, 0.501, 2.0, 30.0] + list(np.geomspace(1e-6, 60.0, 57))
    f1, f2 = [], []
    for z in zs:
        Z = mpmath.mpf(repr(float(z)))
        E = mpmath.exp(-Z)
        f1.append(float(E * (1 + 1 / Z + 1 / Z**2) - 1 / Z**2))
        f2.append(float(3 / Z**2 - E * (1 + 3 / Z + 3 / Z**2)))
    np.savez_compressed(HERE / "brinkman_mp.npz", z=np.array(zs, float), f1=np.array(f1), f2=np.array(f2))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
