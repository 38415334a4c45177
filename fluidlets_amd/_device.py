"""Device memory and streams (PyTorch-ROCm tensors are the allocator; the
compute is the HIP library)."""
from __future__ import annotations

import threading

import numpy as np

_ws_lock = threading.Lock()
_workspaces: dict = {}


def torch():
    import torch as _t
    return _t


def device():
    """The current GPU; raises if there is none (there is no CPU path)."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("fluidlets_amd needs a ROCm GPU (MI355X); no GPU is visible")
    return t.device("cuda", t.cuda.current_device())


def stream_handle(stream=None) -> int:
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(tensor) -> int | None:
    return None if tensor is None else int(tensor.data_ptr())


def soa(points: np.ndarray, dev=None):
    """(N, 3) host points -> (3, N) float64 device tensor."""
    t = torch()
    a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).T)
    return t.from_numpy(a).to(dev or device(), non_blocking=False)


def to_host_aos(u_soa) -> np.ndarray:
    """(3, M) device tensor -> fresh (M, 3) host array."""
    return np.ascontiguousarray(u_soa.cpu().numpy().T)


def eps_arg(eps, n: int, dev=None):
    """Device blob widths and the stride the kernels read them with: a single
    value is passed with stride 0 (broadcast), an (n,) array with stride 1."""
    t = torch()
    e = np.asarray(eps, dtype=np.float64)
    if e.ndim == 0 or e.size == 1:
        return t.full((1,), float(e.reshape(-1)[0]) if e.size else 0.0, dtype=t.float64, device=dev or device()), 0
    e = np.array(np.broadcast_to(e, (n,)), dtype=np.float64)   # a writable copy for torch
    return t.from_numpy(e).to(dev or device()), 1


def workspace(nbytes: int, stream=None):
    """A device byte buffer of at least nbytes, cached per (device, stream).
    Work that shares a stream is ordered, so reuse on one stream is safe."""
    if nbytes <= 0:
        return None
    t = torch()
    key = (t.cuda.current_device(), stream_handle(stream))
    with _ws_lock:
        buf = _workspaces.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = t.empty(int(nbytes), dtype=t.uint8, device=device())
            _workspaces[key] = buf
        return buf
