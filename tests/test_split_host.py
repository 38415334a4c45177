"""The split-bf16 arithmetic of the FP32 LU's bf16x3 trailing products
(fl_sgetrf_ex with FL_LU_BF16X3), restated on the host: the error bounds that
the library's documentation states and its GPU tests rely on."""
import numpy as np


def split_bf16(x):
    """hi, lo of a float32 array: hi = bf16(x), lo = bf16(x - hi), round to
    nearest even (the split the bf16x3 trailing products use)."""
    def bf16(v):
        u = np.ascontiguousarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32)
    hi = bf16(x)
    return hi, bf16((x - hi).astype(np.float32))


def test_bf16_split_bounds_the_product_error():
    """The split keeps 16 bits: |x - hi - lo| <= 2^-17 |x|, and the three-term
    product hi*hi + hi*lo + lo*hi is within 2^-15 |a b| of a b."""
    rng = np.random.default_rng(11)
    a = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    b = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    ah, al = split_bf16(a)
    bh, bl = split_bf16(b)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    assert np.all(np.abs(a64 - ah - al) <= 2.0**-17 * np.abs(a64))
    p3 = ah.astype(np.float64) * bh + ah.astype(np.float64) * bl + al.astype(np.float64) * bh
    assert np.all(np.abs(p3 - a64 * b64) <= 2.0**-15 * np.abs(a64 * b64))
    # the split is exact for values with at most 16 significant bits
    c = np.float32(1.0 + 2.0**-7 + 2.0**-15)
    ch, cl = split_bf16(np.array([c]))
    assert ch[0] + cl[0] == c
