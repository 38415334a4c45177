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


_hip = None
_stream_lock = threading.Lock()
_free_streams: dict = {}       # device -> [ExternalStream] owned here, not leased
_owned_streams: dict = {}      # handle -> ExternalStream (every stream created here)


def _hip_runtime():
    """The HIP runtime that torch has loaded (opened by its soname, so
    dlopen returns the copy already in the process)."""
    global _hip
    if _hip is None:
        import ctypes
        torch()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        _hip.hipStreamCreateWithFlags.restype = ctypes.c_int
        _hip.hipMemcpy2DAsync.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        _hip.hipMemcpy2DAsync.restype = ctypes.c_int
    return _hip


def copy_rows_async(dst: int, dst_pitch: int, src: int, src_pitch: int, width: int, rows: int, kind: str,
                    stream) -> None:
    """`rows` rows of `width` bytes between pitched buffers, one asynchronous
    hipMemcpy2DAsync on `stream`; kind "h2d" or "d2h" (host memory pinned)."""
    if rows <= 0 or width <= 0:
        return
    if width > min(dst_pitch, src_pitch):
        raise ValueError("row width exceeds a pitch")
    rc = _hip_runtime().hipMemcpy2DAsync(dst, dst_pitch, src, src_pitch, width, rows, {"h2d": 1, "d2h": 2}[kind],
                                         stream_handle(stream))
    if rc != 0:
        raise RuntimeError(f"hipMemcpy2DAsync failed with HIP error {rc}")


def lease_stream():
    """A HIP stream for the exclusive use of the caller until release_stream.

    The streams are created here (hipStreamCreateWithFlags, non-blocking,
    default priority) and reused after release; torch's own stream pool is
    not used for work that holds state keyed by its stream, because that pool
    hands out the same 32 streams per device in turn, so two live users could
    end up on one stream."""
    import ctypes
    t = torch()
    dev = t.cuda.current_device()
    with _stream_lock:
        free = _free_streams.setdefault(dev, [])
        if free:
            return free.pop()
    h = ctypes.c_void_p()
    rc = _hip_runtime().hipStreamCreateWithFlags(ctypes.byref(h), 1)   # hipStreamNonBlocking
    if rc != 0:
        raise RuntimeError(f"hipStreamCreateWithFlags failed with HIP error {rc}")
    s = t.cuda.ExternalStream(int(h.value), device=t.device("cuda", dev))
    with _stream_lock:
        _owned_streams[int(h.value)] = s
    return s


def release_stream(s) -> None:
    """Return a stream from lease_stream; work enqueued on it may still run."""
    with _stream_lock:
        if int(s.cuda_stream) not in _owned_streams:
            raise ValueError("release_stream: not a stream from lease_stream")
        _free_streams.setdefault(s.device.index, []).append(s)


def workspace(nbytes: int, stream=None):
    """A device byte buffer of at least nbytes, cached per (device, stream,
    host thread).  The work of one thread on one stream is ordered, so reuse
    there is safe; two threads that share a stream interleave their launches,
    so each gets its own buffer."""
    if nbytes <= 0:
        return None
    t = torch()
    key = (t.cuda.current_device(), stream_handle(stream), threading.get_ident())
    s = stream if stream is not None else t.cuda.current_stream()
    with _ws_lock:
        buf = _workspaces.get(key)
        if buf is None or buf.numel() < nbytes:
            if len(_workspaces) >= 64:
                # drop the buffers of threads that have ended; each was
                # allocated on its own stream, so the allocator reuses it only
                # after the work already queued there
                alive = {th.ident for th in threading.enumerate()}
                for k in [k for k in _workspaces if k[2] not in alive]:
                    del _workspaces[k]
            with t.cuda.stream(s):
                buf = t.empty(int(nbytes), dtype=t.uint8, device=device())
            _workspaces[key] = buf
        return buf
