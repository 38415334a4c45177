import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def kernels_mat():
    return golden("legacy_kernels.npz")


@pytest.fixture(scope="session")
def towers_mat():
    d = golden("legacy_towers.npz")
    n = d["x0"].size // 3
    # legacy arrays are component-major (all x, then y, then z):
    # tests/legacy/test_phase1_port.py:28-30
    d["X"] = d["x0"].reshape(3, n).T.copy()
    d["U"] = d["Vel"].reshape(3, n).T.copy()
    d["F"] = d["Forces"].reshape(3, n).T.copy()
    return d
