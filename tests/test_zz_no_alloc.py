"""fl_eval_noalloc (the kernel sum without a workspace, for callers that
cannot allocate) against fl_eval through the C ABI.

Two calls in a row, the second smaller, so the second runs on scratch that
still holds the first call's partial sums: each must equal fl_eval given a
workspace of the same size bit for bit, and the oracle to the FP64 kernel-sum
bar (1e-12 normwise).  (The rejection of sums that need more scratch than the
fixed array cannot be reached on a 256-CU device: the largest partial-sum
buffer any M, N needs there is about 25 MB of the 32 MiB.)
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _run(entry, kind, M, N, seed, complex_f=False, same=False, ws_bytes=None):
    """One kernel sum through the C ABI; returns (u_re, u_im) on the host."""
    import torch
    from fluidlets_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(seed)
    X = rng.uniform([-1, -1, 0.05], [1, 1, 1.5], (M, 3))
    Y = X if same else rng.uniform([-1, -1, 0.05], [1, 1, 1.5], (N, 3))
    F = rng.standard_normal((N, 3))
    Fi = rng.standard_normal((N, 3)) if complex_f else None
    tgt = _dev(X.T)
    src = tgt if same else _dev(Y.T)
    f = _dev(F.T)
    fi = _dev(Fi.T) if complex_f else None
    eps = _dev(np.array([0.02]))
    u_re = torch.full((3, M), np.nan, dtype=torch.float64, device="cuda")
    u_im = torch.full((3, M), np.nan, dtype=torch.float64, device="cuda") if complex_f else None
    k = _lib.fl_kernel(kind, 0, 3.0 if kind == 2 else 0.0)
    p = lambda t: None if t is None else t.data_ptr()   # noqa: E731
    s = torch.cuda.current_stream().cuda_stream
    if entry == "noalloc":
        st = L.fl_eval_noalloc(k, M, N, p(tgt), p(src), p(eps), 0, p(f), p(fi), 1.0, 0.0, p(u_re), p(u_im),
                               _lib.FL_FP64, s)
    else:
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        st = L.fl_eval(k, M, N, p(tgt), p(src), p(eps), 0, p(f), p(fi), 1.0, 0.0, p(u_re), p(u_im), _lib.FL_FP64,
                       p(ws), ws_bytes, s)
    _lib.check(st, entry)
    torch.cuda.synchronize()
    out = (u_re.cpu().numpy().T.copy(), None if u_im is None else u_im.cpu().numpy().T.copy())
    return out, (X, Y, F, Fi)


CASES = [  # (kernel kind, M, N, complex strengths, targets are the sources)
    (1, 6000, 40000, False, False),     # wall, split sources: partial sums in the scratch
    (2, 1500, 9000, True, False),       # Brinkmanlet(3), complex strengths, smaller: reuses the scratch
    (0, 24000, 24000, False, True),     # targets are the sources (the resistance matvec)
]
NAMES = {0: "stokeslet", 1: "wall", 2: "brinkman:3.0"}


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_equals_fl_eval_twice_in_a_row(order):
    from fluidlets_amd import _lib
    cases = CASES if order == 0 else CASES[::-1]
    for kind, M, N, cf, same in cases:
        print(f"case: fl_eval_noalloc kind {kind} M {M} N {N} complex {cf} same {same}", flush=True)
        assert _lib.lib().fl_eval_workspace_size(M, N, int(cf), _lib.FL_FP64) <= _lib.FL_EVAL_NOALLOC_SCRATCH_BYTES
        (u, ui), (X, Y, F, Fi) = _run("noalloc", kind, M, N, seed=M, complex_f=cf, same=same)
        (v, vi), _ = _run("eval", kind, M, N, seed=M, complex_f=cf, same=same,
                          ws_bytes=_lib.FL_EVAL_NOALLOC_SCRATCH_BYTES)
        assert np.array_equal(u, v)
        if cf:
            assert np.array_equal(ui, vi)
        sub = slice(0, min(M, 700))
        ref = O.evaluate(NAMES[kind], X[sub], Y, F if Fi is None else F + 1j * Fi, 0.02)
        got = u[sub] if ui is None else u[sub] + 1j * ui[sub]
        assert rel_err(got, ref) <= 1e-12

