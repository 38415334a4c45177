"""CPU checks of the C ABI and the host-side logic (no GPU needed)."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "fluidlets_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from fluidlets_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = header_functions()
    assert names, "no declarations found"
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.declared_symbols()) == names
    assert _lib.lib().fl_version() == 1


def test_error_slot_reports_invalid_arguments():
    from fluidlets_amd import _lib
    L = _lib.lib()
    k = _lib.fl_kernel(7, 0, 0.0)
    st = L.fl_eval(k, 1, 1, None, None, None, 0, None, None, 1.0, 0.0, None, None, 0, None, 0, None)
    assert st == _lib.FL_EINVAL
    assert "unknown kernel kind" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(st, "fl_eval")
    k = _lib.fl_kernel(2, 0, -1.0)
    assert L.fl_assemble(k, 1, 1, None, None, None, 0, 1.0, 0.0, None, 3, None, 3, None, None) == _lib.FL_EINVAL
    assert "alpha" in _lib.last_error()


def test_common_normalisation_matches_reference_behaviour():
    from fluidlets_amd.core._common import blob_widths, field_dtype, points
    assert points([1.0, 2.0, 3.0], 3).shape == (1, 3)
    with pytest.raises(ValueError):
        points(np.ones((4, 2)), 3)
    assert blob_widths(0.1, 4).tolist() == [0.1] * 4
    assert field_dtype(1.0, np.ones(3)) == np.float64
    assert field_dtype(1 + 2j, np.ones(3)) == np.complex128
    assert field_dtype(1.0, np.ones(3) * 1j) == np.complex128


def test_kernel_constructors():
    import fluidlets_amd as fl
    assert fl.Stokeslet3D().out_shape == (3, 3) and fl.WallStokeslet3D().dim == 3
    assert fl.Brinkmanlet3D(2).alpha == 2.0
    for bad in (-0.5, 1j):
        with pytest.raises(ValueError):
            fl.Brinkmanlet3D(bad)


def test_nearest_groups_follow_reference_accumulation_order():
    from fluidlets_amd.core.assemble import nearest_groups
    nn = np.array([2, 0, 2, 1, 0, 2])
    order, start = nearest_groups(nn, 4)
    assert start.tolist() == [0, 2, 3, 6, 6]
    assert order.tolist() == [1, 4, 3, 0, 2, 5]
    with pytest.raises(ValueError):
        nearest_groups(np.array([0, 4]), 4)


def test_compute_path_fails_loudly_without_gpu():
    import torch
    import fluidlets_amd as fl
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="GPU"):
        fl.evaluate(fl.Stokeslet3D(), np.ones((2, 3)), np.zeros((2, 3)), np.ones((2, 3)), 0.1)


def test_kernel_plugin_contract_is_explicit():
    """The reference dispatches any Kernel through its `block` function
    (src/fluidlets/kernels/base.py:1-7); the device library implements three
    kernels, so anything else must fail with a clear error rather than an
    AttributeError or a silently substituted built-in kernel."""
    import fluidlets_amd as fl
    from fluidlets_amd import _lib
    from fluidlets_amd.kernels import Kernel, VectorKernel

    assert _lib.kernel_struct(fl.Stokeslet3D()).kind == 0
    assert _lib.kernel_struct(fl.WallStokeslet3D()).kind == 1
    k = _lib.kernel_struct(fl.Brinkmanlet3D(2.5))
    assert (k.kind, k.alpha) == (2, 2.5)

    class Renamed(fl.WallStokeslet3D):          # keeps the built-in block: accepted
        pass
    assert _lib.kernel_struct(Renamed()).kind == 1

    class Custom(VectorKernel):                 # a new kernel defined by its own block
        dim = 3

        @staticmethod
        def block(x, x0, eps, out):
            out[:] = 0.0

    class Overridden(fl.Stokeslet3D):           # a built-in whose block is replaced
        @staticmethod
        def block(x, x0, eps, out):
            out[:] = 1.0

    for bad in (Custom(), Overridden(), Kernel()):
        with pytest.raises(NotImplementedError, match="Stokeslet3D, WallStokeslet3D, Brinkmanlet3D"):
            _lib.kernel_struct(bad)
    with pytest.raises(NotImplementedError):
        fl.evaluate(Custom(), np.ones((2, 3)), np.zeros((2, 3)), np.ones((2, 3)), 0.1)
