"""Fast summation of the kernel sums: a barycentric Lagrange treecode.

The reference evaluates every kernel sum directly (``evaluate``,
src/fluidlets/core/matvec.py:24-49) and defers fast summation
(DESIGN.md:18, 381).  This module adds it beside the direct sum, for N well
beyond what an O(N^2) sum serves: a particle-cluster treecode with
barycentric Lagrange interpolation (Wang, Tlupova & Krasny, SIAM J. Sci.
Comput. 42 (2020) A2993).  It is kernel-independent: the far field of a
cluster of sources is the same regularized kernel summed over (degree+1)^3
proxy sources at the Chebyshev points of the cluster's box, with strengths
interpolated from the cluster's own (``fl_bltc_weights``), so the wall's image
system and the Brinkmanlet need nothing of their own.

The host side here builds the trees and the interaction lists (NumPy):

- sources (and targets) are ordered by a binary tree that splits each box at
  the median along its longest side until a leaf holds ``leaf_size`` points;
  every node is a contiguous range of the ordered points with a tight box;
- a target leaf B and a source node C interact through C's proxies when
  (r_B + r_C) < theta |c_B - c_C| (box half-diagonals r, centres c) and C
  holds more sources than proxies; directly over C's whole range when the
  criterion holds but C is small, or C is a leaf; otherwise through C's two
  children.  Every (target leaf, source) pair is covered exactly once.

The device side (``fl_bltc_eval``) walks each block of 256 targets' list of
segments with the direct sum's pair code, in list order (deterministic).

One blob width (scalar eps) is required: the proxies carry it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _device as D
from ._lib import check, kernel_struct, lib
from .core._common import blob_widths, field_dtype, points

BLOCK = 256   # targets per workgroup of the evaluation kernel (one per lane)


@dataclass
class Tree:
    """A binary tree over points: node k covers order[begin[k]:end[k]]."""
    order: np.ndarray      # (N,) int64, the points' tree order
    begin: np.ndarray      # (K,) int64
    end: np.ndarray        # (K,) int64
    lo: np.ndarray         # (K, 3) tight box
    hi: np.ndarray         # (K, 3)
    child: np.ndarray      # (K, 2) int64, -1 for a leaf

    @property
    def centre(self):
        return 0.5 * (self.lo + self.hi)

    @property
    def radius(self):
        return 0.5 * np.linalg.norm(self.hi - self.lo, axis=1)

    @property
    def leaves(self):
        return np.flatnonzero(self.child[:, 0] < 0)


def build_tree(X: np.ndarray, leaf_size: int) -> Tree:
    """Median splits along the longest side of each node's tight box until a
    node holds at most ``leaf_size`` points (breadth first, so a node's
    children follow it and each level's nodes are contiguous)."""
    X = np.asarray(X, dtype=np.float64)
    N = X.shape[0]
    if leaf_size < 1:
        raise ValueError("leaf_size must be >= 1")
    order = np.arange(N, dtype=np.int64)
    begin, end, child = [0], [N], []
    lo, hi = [], []
    k = 0
    while k < len(begin):
        b, e = begin[k], end[k]
        pts = X[order[b:e]]
        l, h = (pts.min(axis=0), pts.max(axis=0)) if e > b else (np.zeros(3), np.zeros(3))
        lo.append(l)
        hi.append(h)
        if e - b > leaf_size:
            d = int(np.argmax(h - l))
            m = (b + e) // 2
            idx = order[b:e]
            order[b:e] = idx[np.argpartition(X[idx, d], m - b)]
            child.append([len(begin), len(begin) + 1])
            begin += [b, m]
            end += [m, e]
        else:
            child.append([-1, -1])
        k += 1
    return Tree(order, np.array(begin, dtype=np.int64), np.array(end, dtype=np.int64), np.array(lo),
                np.array(hi), np.array(child, dtype=np.int64))


def interaction_lists(tt: Tree, st: Tree, theta: float, nproxy: int):
    """For each leaf of the target tree, the source nodes it takes directly
    (their whole range) and those it takes through proxies.  Returns
    (direct, approx): lists of (target leaf, source node) index arrays."""
    if not 0.0 <= theta < 1.0:
        raise ValueError("theta must be in [0, 1)")
    tl = tt.leaves
    B = tl.copy()
    C = np.zeros_like(B)
    cB, rB = tt.centre, tt.radius
    cC, rC = st.centre, st.radius
    size = st.end - st.begin
    dirB, dirC, appB, appC = [], [], [], []
    while B.size:
        dist = np.linalg.norm(cB[B] - cC[C], axis=1)
        mac = (rB[B] + rC[C]) < theta * dist
        leaf = st.child[C, 0] < 0
        app = mac & (size[C] > nproxy)
        dr = ~app & (mac | leaf)
        ex = ~app & ~dr
        appB.append(B[app]); appC.append(C[app])
        dirB.append(B[dr]); dirC.append(C[dr])
        B = np.repeat(B[ex], 2)
        C = st.child[C[ex]].reshape(-1)
    cat = lambda a: np.concatenate(a) if a else np.zeros(0, np.int64)
    return (cat(dirB), cat(dirC)), (cat(appB), cat(appC))


def chunk_table(begin: np.ndarray, end: np.ndarray, target: int = 16384):
    """Splits each cluster's source range [begin, end) into chunks of at most
    ``size`` sources, ``size`` = 4096 doubled until there are no more than
    ``target`` chunks or every cluster is one chunk (enough workgroups to
    fill the GPU several times, a bounded partial-strength workspace).
    Returns per cluster the chunk count and first chunk, and the chunk rows
    (cluster, begin, end)."""
    begin = np.asarray(begin, dtype=np.int64)
    end = np.asarray(end, dtype=np.int64)
    n = end - begin
    size = 4096
    while np.sum(-(-n // size)) > target and size < n.max():
        size *= 2
    nch = -(-n // size)
    first = (np.cumsum(nch) - nch).astype(np.int64)
    cl = np.repeat(np.arange(n.size, dtype=np.int64), nch)
    k = np.arange(cl.size, dtype=np.int64) - first[cl]
    b = begin[cl] + k * size
    chunk = np.stack([cl, b, np.minimum(b + size, end[cl])], axis=1).astype(np.int64)
    return nch.astype(np.int64), first, chunk.reshape(-1, 3)


class Treecode:
    """A treecode for fixed targets, sources and blob width: ``apply`` sums
    the kernel over new strengths with the same trees and lists (the GMRES
    matvec's use), ``evaluate_treecode`` builds one for a single sum."""

    def __init__(self, kernel, targets, sources, eps, degree: int = 8, theta: float = 0.7,
                 leaf_size: int = 1024, device=None):
        kernel_struct(kernel)
        if not 1 <= int(degree) <= 12:
            raise ValueError("degree must be 1 .. 12")
        self.kernel = kernel
        self.degree, self.theta, self.leaf_size = int(degree), float(theta), int(leaf_size)
        T = points(targets, kernel.dim)
        S = points(sources, kernel.dim)
        self.M, self.N = T.shape[0], S.shape[0]
        e = blob_widths(eps, max(self.N, 1))
        if self.N and not np.all(e == e[0]):
            raise ValueError("the treecode needs one blob width (the proxy sources carry it)")
        self.eps = float(e[0]) if self.N else 1.0
        nq = (self.degree + 1) ** 3
        self.nproxy = nq
        self.same = T.shape == S.shape and np.array_equal(T, S)
        self.st = build_tree(S, self.leaf_size)
        self.tt = self.st if self.same else build_tree(T, self.leaf_size)
        (dB, dC), (aB, aC) = interaction_lists(self.tt, self.st, self.theta, nq)
        self.n_direct_pairs = int(np.sum((self.st.end - self.st.begin)[dC] * (self.tt.end - self.tt.begin)[dB]))
        self.n_proxy_pairs = int(np.sum((self.tt.end - self.tt.begin)[aB]) * nq)
        # clusters with proxies, their first proxy and their chunks (the
        # moments kernel sums each chunk in a workgroup of its own)
        pc = np.unique(aC)
        self.P = int(pc.size * nq)
        base = np.full(self.st.begin.size, -1, np.int64)
        base[pc] = np.arange(pc.size, dtype=np.int64) * nq
        cb, ce = self.st.begin[pc], self.st.end[pc]
        nch, first, chunk = chunk_table(cb, ce)
        cl = np.stack([cb, ce, base[pc], first, nch], axis=1).astype(np.int64)
        box = np.concatenate([self.st.lo[pc], self.st.hi[pc]], axis=1)
        # segments per target leaf: direct runs (merged where contiguous), then proxies
        leaves = self.tt.leaves
        seg_rows, blk_rows = [], []
        order_d = np.lexsort((self.st.begin[dC], dB))
        dB, dC = dB[order_d], dC[order_d]
        order_a = np.lexsort((aC, aB))
        aB, aC = aB[order_a], aC[order_a]
        d_start = np.searchsorted(dB, leaves, side="left")
        d_stop = np.searchsorted(dB, leaves, side="right")
        a_start = np.searchsorted(aB, leaves, side="left")
        a_stop = np.searchsorted(aB, leaves, side="right")
        nseg = 0
        for i, lf in enumerate(leaves):
            segs = []
            for c in dC[d_start[i]:d_stop[i]]:
                b, e = int(self.st.begin[c]), int(self.st.end[c])
                if e == b:
                    continue
                if segs and segs[-1][0] == 0 and segs[-1][1] + segs[-1][2] == b:
                    segs[-1][2] += e - b
                else:
                    segs.append([0, b, e - b])
            for c in aC[a_start[i]:a_stop[i]]:
                segs.append([1, int(base[c]), nq])
            s0 = nseg
            seg_rows.extend(segs)
            nseg += len(segs)
            tb, te = int(self.tt.begin[lf]), int(self.tt.end[lf])
            for t0 in range(tb, te, BLOCK):
                blk_rows.append([t0, min(BLOCK, te - t0), s0, nseg])
        self.nblk = len(blk_rows)
        self.nseg = nseg
        dev = device if device is not None else D.device()
        t = D.torch()
        as64 = lambda a: t.from_numpy(np.ascontiguousarray(a, dtype=np.int64).reshape(-1)).to(dev)
        self._cl = as64(cl)
        self.nchunk = int(chunk.shape[0])
        self._chunk = as64(chunk)
        self._ws = None
        self._box = t.from_numpy(np.ascontiguousarray(box, dtype=np.float64).reshape(-1)).to(dev)
        self._blk = as64(np.array(blk_rows, dtype=np.int64).reshape(-1, 4))
        self._seg = as64(np.array(seg_rows, dtype=np.int64).reshape(-1, 3))
        self._tidx = as64(self.tt.order)
        self._src = D.soa(S[self.st.order], dev)
        self._tgt = self._src if self.same else D.soa(T[self.tt.order], dev)
        self._px = t.empty((3, max(self.P, 1)), dtype=t.float64, device=dev)
        self._pf = t.empty((3, max(self.P, 1)), dtype=t.float64, device=dev)
        self._pfi = None
        self.ncl = int(pc.size)

    def apply_device(self, f_re, f_im=None, mu=1.0, stream=None):
        """Kernel sums for strengths f (3, N) device tensors in the sources'
        original order; returns (u_re, u_im) (3, M) device tensors (u_im None
        for a real result), enqueued on ``stream``."""
        t = D.torch()
        mu = complex(mu)
        L = lib()
        dev = self._src.device
        order = t.from_numpy(self.st.order).to(dev)
        fr = f_re[:, order].contiguous()
        fi = f_im[:, order].contiguous() if f_im is not None else None
        complex_out = fi is not None or mu.imag != 0.0
        u_re = t.empty((3, self.M), dtype=t.float64, device=dev)
        u_im = t.empty((3, self.M), dtype=t.float64, device=dev) if complex_out else None
        if fi is not None and self._pfi is None:
            self._pfi = t.empty_like(self._pf)
        pfi = self._pfi if fi is not None else None
        nws = max(self.nchunk, 1) * 3 * (2 if fi is not None else 1) * self.nproxy
        if self._ws is None or self._ws.numel() < nws:
            self._ws = t.empty(nws, dtype=t.float64, device=dev)
        sh = D.stream_handle(stream)
        check(L.fl_bltc_weights(self.ncl, D.ptr(self._cl), D.ptr(self._box), self.degree, self.nchunk,
                                D.ptr(self._chunk), self.N, D.ptr(self._src), D.ptr(fr), D.ptr(fi), self.P,
                                D.ptr(self._px), D.ptr(self._pf), D.ptr(pfi), D.ptr(self._ws), sh),
              "fl_bltc_weights")
        k = kernel_struct(self.kernel)
        check(L.fl_bltc_eval(k, self.M, D.ptr(self._tgt), D.ptr(self._tidx), self.nblk, D.ptr(self._blk),
                             D.ptr(self._seg), self.N, D.ptr(self._src), D.ptr(fr), D.ptr(fi), self.P,
                             D.ptr(self._px), D.ptr(self._pf), D.ptr(pfi), self.eps, mu.real, mu.imag,
                             D.ptr(u_re), D.ptr(u_im) if complex_out else None, sh), "fl_bltc_eval")
        return u_re, u_im


def evaluate_treecode(kernel, targets, sources, strengths, eps, mu=1.0, degree: int = 8, theta: float = 0.7,
                      leaf_size: int = 1024) -> np.ndarray:
    """``evaluate`` (the reference's core/matvec.py:24-49 contract: shapes,
    dtypes, complex promotion) by the treecode: the far field of each source
    cluster through (degree+1)^3 proxies where (r_B + r_C) < theta |c_B - c_C|.
    One blob width only.  Returns (M, 3)."""
    kernel_struct(kernel)
    T = points(targets, kernel.dim)
    S = points(sources, kernel.dim)
    N = S.shape[0]
    dtype = field_dtype(mu, strengths)
    f = np.ascontiguousarray(np.asarray(strengths, dtype=dtype).reshape(N, 3))
    M = T.shape[0]
    if M == 0 or N == 0:
        return np.zeros((M, 3), dtype=dtype) / mu
    tc = Treecode(kernel, T, S, eps, degree=degree, theta=theta, leaf_size=leaf_size)
    dev = tc._src.device
    if np.iscomplexobj(f) and np.any(f.imag != 0):
        f_re, f_im = D.soa(np.ascontiguousarray(f.real), dev), D.soa(np.ascontiguousarray(f.imag), dev)
    else:
        f_re, f_im = D.soa(np.ascontiguousarray(f.real if np.iscomplexobj(f) else f), dev), None
    u_re, u_im = tc.apply_device(f_re, f_im, mu)
    out = D.to_host_aos(u_re)
    if np.issubdtype(dtype, np.complexfloating):
        out = out + 1j * D.to_host_aos(u_im) if u_im is not None else out.astype(dtype)
    return out
