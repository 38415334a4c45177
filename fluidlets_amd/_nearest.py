"""Nearest force point of each quadrature point, on the device (fl_nearest)."""
from __future__ import annotations

import numpy as np

from . import _device as D
from ._lib import check, lib


def nearest_device(quad_soa, force_soa, out=None, stream=None):
    t = D.torch()
    Q, N = int(quad_soa.shape[1]), int(force_soa.shape[1])
    if out is None:
        out = t.empty((Q,), dtype=t.int64, device=quad_soa.device)
    check(lib().fl_nearest(Q, N, D.ptr(quad_soa), D.ptr(force_soa), D.ptr(out), D.stream_handle(stream)),
          "fl_nearest")
    return out


def nearest_neighbors(quad_points, force_points) -> np.ndarray:
    """Drop-in for src/fluidlets/core/assemble.py:61-65: (Q,) int64 indices."""
    quad = np.atleast_2d(np.asarray(quad_points, dtype=np.float64))
    force = np.atleast_2d(np.asarray(force_points, dtype=np.float64))
    if quad.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    if force.shape[0] == 0:
        raise ValueError("no force points")
    dev = D.device()
    return nearest_device(D.soa(quad, dev), D.soa(force, dev)).cpu().numpy()
