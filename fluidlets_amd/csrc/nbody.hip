// Tiled N-body kernel sums of the regularized Stokeslet family on gfx950.
//
// One target per lane (TPL targets per lane when register-blocked), 256-lane
// workgroups, source tiles of 256 staged in LDS from coalesced
// structure-of-arrays loads and read back as wave-wide broadcasts.  Every lane
// walks the sources in index order, so each target's sum has the same order as
// the reference loop (src/fluidlets/core/matvec.py:15-21); when the source
// range is split across workgroups (to fill 256 CUs at small M), each chunk
// keeps that order and the chunk sums are combined in chunk order by a second
// kernel, so results are deterministic.
//
// Pair formulas (mu = 1; constant factors 1/(8 pi) or 1/(4 pi) applied once
// per target in the epilogue):
//   Stokeslet   kernels/stokeslet3d.py:22-50
//   Wall        kernels/images.py:26-89, free + image fused in one pass that
//               shares d0, d1, d0^2 + d1^2 and the source tile
//   Brinkmanlet kernels/brinkmanlet3d.py:52-104, with 1/z = rsqrt(Q)/alpha,
//               e^{-z} from a table of 2^(j/64) in LDS and a degree-5
//               polynomial, and the 24-term series below z = 0.5 in FP64
#include "fl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace fl {
namespace {

constexpr int WG = 256;
constexpr int TILE = 256;
constexpr int NTERMS = 24;

struct SeriesCoef {
    double c1[NTERMS];
    double c2[NTERMS];
};

// c1_n = (-1)^n [1/(n+2)! - 1/(n+1)! + 1/n!],
// c2_n = (-1)^(n+1) [3/(n+2)! - 3/(n+1)! + 1/n!]  (Taylor coefficients of
// e^{-z}(1 + 1/z + 1/z^2) - 1/z^2 and 3/z^2 - e^{-z}(1 + 3/z + 3/z^2)).
constexpr SeriesCoef make_series()
{
    SeriesCoef c{};
    double fact[NTERMS + 3] = {};
    fact[0] = 1.0;
    for (int k = 1; k < NTERMS + 3; ++k) fact[k] = fact[k - 1] * k;
    for (int n = 0; n < NTERMS; ++n) {
        double s = (n % 2 == 0) ? 1.0 : -1.0;
        c.c1[n] = s * (1.0 / fact[n + 2] - 1.0 / fact[n + 1] + 1.0 / fact[n]);
        c.c2[n] = -s * (3.0 / fact[n + 2] - 3.0 / fact[n + 1] + 1.0 / fact[n]);
    }
    return c;
}

__constant__ SeriesCoef kSeries = make_series();

// 1/sqrt(x) for finite x > 0 (here x = r^2 + eps^2 >= eps^2 > 0): the
// hardware estimate y0 (v_rsq_f64) refined by one second-order step of the
// expansion (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + ..., e = 1 - x y0^2, which
// takes the estimate's relative error d to O(d^3).  No special-value handling.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    double y0 = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// 2^(j/128), j = 0..127, each the double nearest the exact value (computed to
// 60 digits and rounded once).
__constant__ double kExp2Frac[128] = {
    0x1.0000000000000p+0  , 0x1.0163da9fb3335p+0  , 0x1.02c9a3e778061p+0  , 0x1.04315e86e7f85p+0,
    0x1.059b0d3158574p+0  , 0x1.0706b29ddf6dep+0  , 0x1.0874518759bc8p+0  , 0x1.09e3ecac6f383p+0,
    0x1.0b5586cf9890fp+0  , 0x1.0cc922b7247f7p+0  , 0x1.0e3ec32d3d1a2p+0  , 0x1.0fb66affed31bp+0,
    0x1.11301d0125b51p+0  , 0x1.12abdc06c31ccp+0  , 0x1.1429aaea92de0p+0  , 0x1.15a98c8a58e51p+0,
    0x1.172b83c7d517bp+0  , 0x1.18af9388c8deap+0  , 0x1.1a35beb6fcb75p+0  , 0x1.1bbe084045cd4p+0,
    0x1.1d4873168b9aap+0  , 0x1.1ed5022fcd91dp+0  , 0x1.2063b88628cd6p+0  , 0x1.21f49917ddc96p+0,
    0x1.2387a6e756238p+0  , 0x1.251ce4fb2a63fp+0  , 0x1.26b4565e27cddp+0  , 0x1.284dfe1f56381p+0,
    0x1.29e9df51fdee1p+0  , 0x1.2b87fd0dad990p+0  , 0x1.2d285a6e4030bp+0  , 0x1.2ecafa93e2f56p+0,
    0x1.306fe0a31b715p+0  , 0x1.32170fc4cd831p+0  , 0x1.33c08b26416ffp+0  , 0x1.356c55f929ff1p+0,
    0x1.371a7373aa9cbp+0  , 0x1.38cae6d05d866p+0  , 0x1.3a7db34e59ff7p+0  , 0x1.3c32dc313a8e5p+0,
    0x1.3dea64c123422p+0  , 0x1.3fa4504ac801cp+0  , 0x1.4160a21f72e2ap+0  , 0x1.431f5d950a897p+0,
    0x1.44e086061892dp+0  , 0x1.46a41ed1d0057p+0  , 0x1.486a2b5c13cd0p+0  , 0x1.4a32af0d7d3dep+0,
    0x1.4bfdad5362a27p+0  , 0x1.4dcb299fddd0dp+0  , 0x1.4f9b2769d2ca7p+0  , 0x1.516daa2cf6642p+0,
    0x1.5342b569d4f82p+0  , 0x1.551a4ca5d920fp+0  , 0x1.56f4736b527dap+0  , 0x1.58d12d497c7fdp+0,
    0x1.5ab07dd485429p+0  , 0x1.5c9268a5946b7p+0  , 0x1.5e76f15ad2148p+0  , 0x1.605e1b976dc09p+0,
    0x1.6247eb03a5585p+0  , 0x1.6434634ccc320p+0  , 0x1.6623882552225p+0  , 0x1.68155d44ca973p+0,
    0x1.6a09e667f3bcdp+0  , 0x1.6c012750bdabfp+0  , 0x1.6dfb23c651a2fp+0  , 0x1.6ff7df9519484p+0,
    0x1.71f75e8ec5f74p+0  , 0x1.73f9a48a58174p+0  , 0x1.75feb564267c9p+0  , 0x1.780694fde5d3fp+0,
    0x1.7a11473eb0187p+0  , 0x1.7c1ed0130c132p+0  , 0x1.7e2f336cf4e62p+0  , 0x1.80427543e1a12p+0,
    0x1.82589994cce13p+0  , 0x1.8471a4623c7adp+0  , 0x1.868d99b4492edp+0  , 0x1.88ac7d98a6699p+0,
    0x1.8ace5422aa0dbp+0  , 0x1.8cf3216b5448cp+0  , 0x1.8f1ae99157736p+0  , 0x1.9145b0b91ffc6p+0,
    0x1.93737b0cdc5e5p+0  , 0x1.95a44cbc8520fp+0  , 0x1.97d829fde4e50p+0  , 0x1.9a0f170ca07bap+0,
    0x1.9c49182a3f090p+0  , 0x1.9e86319e32323p+0  , 0x1.a0c667b5de565p+0  , 0x1.a309bec4a2d33p+0,
    0x1.a5503b23e255dp+0  , 0x1.a799e1330b358p+0  , 0x1.a9e6b5579fdbfp+0  , 0x1.ac36bbfd3f37ap+0,
    0x1.ae89f995ad3adp+0  , 0x1.b0e07298db666p+0  , 0x1.b33a2b84f15fbp+0  , 0x1.b59728de5593ap+0,
    0x1.b7f76f2fb5e47p+0  , 0x1.ba5b030a1064ap+0  , 0x1.bcc1e904bc1d2p+0  , 0x1.bf2c25bd71e09p+0,
    0x1.c199bdd85529cp+0  , 0x1.c40ab5fffd07ap+0  , 0x1.c67f12e57d14bp+0  , 0x1.c8f6d9406e7b5p+0,
    0x1.cb720dcef9069p+0  , 0x1.cdf0b555dc3fap+0  , 0x1.d072d4a07897cp+0  , 0x1.d2f87080d89f2p+0,
    0x1.d5818dcfba487p+0  , 0x1.d80e316c98398p+0  , 0x1.da9e603db3285p+0  , 0x1.dd321f301b460p+0,
    0x1.dfc97337b9b5fp+0  , 0x1.e264614f5a129p+0  , 0x1.e502ee78b3ff6p+0  , 0x1.e7a51fbc74c83p+0,
    0x1.ea4afa2a490dap+0  , 0x1.ecf482d8e67f1p+0  , 0x1.efa1bee615a27p+0  , 0x1.f252b376bba97p+0,
    0x1.f50765b6e4540p+0  , 0x1.f7bfdad9cbe14p+0  , 0x1.fa7c1819e90d8p+0  , 0x1.fd3c22b8f71f1p+0,
};

// LDS of a Brinkmanlet kernel: the table of exp_neg (128 doubles; the lanes
// of a wave index it at different j, which LDS serves in one pass) and the
// series coefficients c1, c2 (read only on the rare near-pair path, so they
// hold no registers across the pair loop).
constexpr int BRINK_TAB = 128;
constexpr int BRINK_LDS = BRINK_TAB + 2 * NTERMS + 1;
constexpr int BRINK_C3 = BRINK_TAB + 2 * NTERMS;   // c3 of the exp polynomial, kept in a vector register

__device__ __forceinline__ void stage_brinkman(double* bl)
{
    const int i = threadIdx.x;
    if (i < BRINK_TAB) bl[i] = kExp2Frac[i];
    if (i < NTERMS) {
        bl[BRINK_TAB + i] = kSeries.c1[i];
        bl[BRINK_TAB + NTERMS + i] = kSeries.c2[i];
    }
    if (i == 0) bl[BRINK_C3] = 0x1.55555b7bae95fp-3;
}

// a * b + c, one v_fma_f64 with b from a scalar and c from a vector register
// (the compiler otherwise spends a register copy on each fused multiply-add
// whose loop-invariant addend it would overwrite).
__device__ __forceinline__ double fma_sv(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}

// e^{-z} for z >= 0, from the LDS block etab of a Brinkmanlet kernel
// (etab[j] = 2^(j/128), etab[BRINK_C3] = c3): with
// m = rint(-128 z / ln 2) = 128 n + j (0 <= j < 128, n = floor(m / 128)),
//   e^{-z} = 2^n * 2^(j/128) * e^r,   r = -z - m ln2/128,  |r| <= ln2/256,
// where m ln2/128 is taken with ln2/128 split in a high and a low part (the
// two fused multiply-adds leave r exact to about 1e-19), and e^r is
//   1 + r + r^2 / 2 + c3 r^3 + c4 r^4,
// c3 and c4 fitted at the three Chebyshev nodes of [-ln2/256, ln2/256] for
// (e^r - 1 - r) / r^2 (relative error 3.0e-16 over the interval, against
// 1.2e-15 for the Taylor coefficients 1/6, 1/24; the 64-entry table needed
// degree 5, one more fused multiply-add).  The result is within about 2 ulp.
// m is rounded by adding 1.5 * 2^52 (the
// sum's low word is then m as a two's-complement integer).  z is clamped at
// 1000, where e^{-z} underflows to 0 anyway.
__device__ __forceinline__ double exp_neg(double z, const double* etab)
{
    constexpr double kScale = -184.6649652337873;      // -128 / ln 2
    constexpr double kLhi = 0.0054152123481245725;      // ln 2 / 128, high part
    constexpr double kLlo = 1.8117553233174215e-19;     // ln 2 / 128 - kLhi
    constexpr double kShift = 6755399441055744.0;       // 1.5 * 2^52
    asm("v_min_f64 %0, %1, %2" : "=v"(z) : "v"(z), "s"(1000.0));   // z is never NaN here
    const double ms = fma(z, kScale, kShift);
    const double m = ms - kShift;
    const int mi = __double2loint(ms);
    double r = fma(m, -kLhi, -z);
    r = fma(m, -kLlo, r);
    const double T = etab[mi & (BRINK_TAB - 1)];
    double p = fma_sv(r, 0x1.5555596ee628dp-5, etab[BRINK_C3]);   // c4 r + c3
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(T * p, mi >> 7);
}

// ---------------------------------------------------------------- FP64 pairs
// acc[3r + a] += (G f_r)_a in units of the per-kernel scale.

template <int NF>
__device__ __forceinline__ void pair_stokeslet(double d0, double d1, double d2, double e2, const double* f,
                                               double* acc)
{
    double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    double Q = r2 + e2;
    double rs = rsqrt_pos(Q);
    double s3 = rs * rs * rs;
    double h1 = (Q + e2) * s3;  // (r^2 + 2 eps^2) / Q^{3/2}
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        double fd = f[3 * r] * d0 + f[3 * r + 1] * d1 + f[3 * r + 2] * d2;
        double t = s3 * fd;
        acc[3 * r] += h1 * f[3 * r] + t * d0;
        acc[3 * r + 1] += h1 * f[3 * r + 1] + t * d1;
        acc[3 * r + 2] += h1 * f[3 * r + 2] + t * d2;
    }
}

// Free Stokeslet + no-slip image system (kernels/images.py:26-104) as one
// 3x3 block, in units of 1/(8 pi).  For a target at height z and a source at
// height h (dzf = z - h, dzi = z + h, in-plane displacement d0, d1), summing
// the free block and the image block of images.py:48-74 (image coefficients
// doubled to share the 1/(8 pi) scale) and collecting terms gives
//   G = [[C + b d0^2,  b d0 d1,     g d0],
//        [b d0 d1,     C + b d1^2,  g d1],
//        [t d0,        t d1,        Z   ]]
// with Qf = r_f^2 + e2, Qi = r_i^2 + e2, s3f = Qf^-3/2, s3 = Qi^-3/2,
// s5 = Qi^-5/2, h1f = (Qf + e2) s3f, a = s3f - s3, k = 6 h z s5:
//   b = a + k,  g = dzf a - k dzi,  t = dzf a + k dzi,
//   C = h1f - Qi^-1/2 - s3 (e2 + 2 h z) + e2 k,
//   Z = h1f + dzf^2 s3f - Qi^-1/2 - s3 (e2 + h^2 + z^2) - k (e2 + dzi^2)
//     = C + dzf^2 a - k (2 e2 + dzi^2)          (dzf^2 + 2 h z = h^2 + z^2).
// (Swapping target and source maps G to its transpose: g <-> -t under
// d -> -d, the reciprocal symmetry of the image system.)
struct WallBlock {
    double C, b, g, t, Z;
};

// h2 = 2h, h6 = 6h, e2x2 = 2 e2 (per source).
__device__ __forceinline__ WallBlock wall_block(double d0, double d1, double z, double h, double e2, double h2,
                                                double h6, double e2x2)
{
    const double dzf = z - h, dzi = z + h;
    const double tq = fma(d0, d0, d1 * d1) + e2;
    const double Qf = fma(dzf, dzf, tq);
    const double Qi = fma(dzi, dzi, tq);
    const double rf = rsqrt_pos(Qf);
    const double ri = rsqrt_pos(Qi);
    const double s3f = rf * rf * rf;
    const double ri2 = ri * ri;
    const double s3 = ri2 * ri;
    const double s5 = s3 * ri2;
    const double hr = (Qf + e2) * s3f - ri;
    const double k = (h6 * z) * s5;
    const double a = s3f - s3;
    const double kd = k * dzi;
    const double p = dzf * a;
    WallBlock B;
    B.b = a + k;
    B.g = p - kd;
    B.t = p + kd;
    B.C = fma(e2, k, fma(-s3, fma(h2, z, e2), hr));
    B.Z = fma(-k, fma(dzi, dzi, e2x2), fma(dzf, p, B.C));
    return B;
}

template <int NF>
__device__ __forceinline__ void pair_wall(double d0, double d1, double z, double h, double e2, double h2, double h6,
                                          double e2x2, const double* f, double* acc)
{
    const WallBlock B = wall_block(d0, d1, z, h, e2, h2, h6, e2x2);
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        const double f0 = f[3 * r], f1 = f[3 * r + 1], f2 = f[3 * r + 2];
        const double w = fma(d0, f0, d1 * f1);
        const double v = fma(B.b, w, B.g * f2);
        acc[3 * r] = fma(B.C, f0, fma(d0, v, acc[3 * r]));
        acc[3 * r + 1] = fma(B.C, f1, fma(d1, v, acc[3 * r + 1]));
        acc[3 * r + 2] = fma(B.t, w, fma(B.Z, f2, acc[3 * r + 2]));
    }
}

// Brinkmanlet block G = H1 I + H2 d d^T (kernels/brinkmanlet3d.py:52-104):
// with z = alpha sqrt(Q), Q = r^2 + e2, rs = Q^{-1/2},
//   H2 = b2(z) rs^3,  H1 = b1(z) rs + e2 H2,
//   b1 = e^{-z}(1 + 1/z + 1/z^2) - 1/z^2,  b2 = 3/z^2 - e^{-z}(1 + 3/z + 3/z^2)
//      = 2 e^{-z} - 3 b1.
// The FP64 kernels work in coordinates scaled by alpha (x' = alpha x,
// e2' = alpha^2 e2; alpha > 0): then Q' = alpha^2 Q, u = Q'^{-1/2} = 1/z,
// z = Q' u, and G f = alpha [H1' f + H2' (d'.f) d'] with
//   H2' = b2 u^3,  H1' = b1 u + e2' H2',
// the factor alpha applied once per target with the kernel's scale.
// brinkman_radial evaluates the closed form for every pair (straight-line
// code, so the pairs of several targets interleave) and returns z; pairs with
// z < 0.5, where the closed form cancels, are then redone by
// brinkman_radial_near from the 24-term series (a branch that is rarely taken:
// such pairs are closer than 0.5 / alpha).  bl is the kernel's LDS block.
__device__ __forceinline__ double brinkman_radial(double Q, double u, double e2, const double* bl, double& H1,
                                                  double& H2)
{
    const double z = Q * u;
    const double E = exp_neg(z, bl);
    const double v = u * u;
    const double b1 = fma(E, fma(u, u, u + 1.0), -v);
    const double b2 = fma(-3.0, b1, E + E);
    H2 = b2 * (u * v);
    H1 = fma(b1, u, e2 * H2);
    return z;
}

// b1(z), b2(z) from their Taylor series with the coefficients in LDS.
__device__ __forceinline__ void brinkman_series_lds(double z, const double* bl, double& b1, double& b2)
{
    int k = 0;
    asm volatile("" : "+s"(k));   // the coefficient reads stay inside the calling branch
    const double* c1 = bl + BRINK_TAB + k;
    const double* c2 = c1 + NTERMS;
    b1 = 0.0;
    b2 = 0.0;
#pragma unroll
    for (int n = NTERMS - 1; n >= 0; --n) {
        b1 = fma(b1, z, c1[n]);
        b2 = fma(b2, z, c2[n]);
    }
}

__device__ __forceinline__ void brinkman_radial_near(double z, double u, double e2, const double* bl, double& H1,
                                                     double& H2)
{
    double b1, b2;
    brinkman_series_lds(z, bl, b1, b2);
    H2 = b2 * (u * u * u);
    H1 = fma(b1, u, e2 * H2);
}

// The Brinkmanlet pairs of TPL targets and one source (strengths f, NF
// sets), all in alpha-scaled coordinates.
template <int NF, int TPL>
__device__ __forceinline__ void pairs_brinkman(const double* tx, const double* ty, const double* tz, const double* r,
                                               const double* bl, const double* f, double (*acc)[3 * NF])
{
    const double e2 = r[3];
    double d[TPL][3], H1[TPL], H2[TPL], z[TPL], rs[TPL];
    bool near = false;
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        d[t][0] = tx[t] - r[0];
        d[t][1] = ty[t] - r[1];
        d[t][2] = tz[t] - r[2];
        const double Q = fma(d[t][0], d[t][0], fma(d[t][1], d[t][1], fma(d[t][2], d[t][2], e2)));
        rs[t] = rsqrt_pos(Q);
        z[t] = brinkman_radial(Q, rs[t], e2, bl, H1[t], H2[t]);
        near |= z[t] < 0.5;
    }
    if (near) {
#pragma unroll
        for (int t = 0; t < TPL; ++t)
            if (z[t] < 0.5) brinkman_radial_near(z[t], rs[t], e2, bl, H1[t], H2[t]);
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            const double* fq = f + 3 * q;
            const double u = H2[t] * fma(d[t][0], fq[0], fma(d[t][1], fq[1], d[t][2] * fq[2]));
            acc[t][3 * q] = fma(H1[t], fq[0], fma(u, d[t][0], acc[t][3 * q]));
            acc[t][3 * q + 1] = fma(H1[t], fq[1], fma(u, d[t][1], acc[t][3 * q + 1]));
            acc[t][3 * q + 2] = fma(H1[t], fq[2], fma(u, d[t][2], acc[t][3 * q + 2]));
        }
    }
}

// ---------------------------------------------------------------- FP32 pairs
// Displacements are formed in FP64 and rounded once; the Brinkmanlet series
// branch (z < 0.5, where the closed form cancels) runs in FP64.

template <int NF>
__device__ __forceinline__ void pair_stokeslet32(float d0, float d1, float d2, float e2, const float* f, float* acc)
{
    float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    float Q = r2 + e2;
    float rs = rsqrtf(Q);
    float s3 = rs * rs * rs;
    float h1 = (Q + e2) * s3;
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        float fd = f[3 * r] * d0 + f[3 * r + 1] * d1 + f[3 * r + 2] * d2;
        float t = s3 * fd;
        acc[3 * r] += h1 * f[3 * r] + t * d0;
        acc[3 * r + 1] += h1 * f[3 * r + 1] + t * d1;
        acc[3 * r + 2] += h1 * f[3 * r + 2] + t * d2;
    }
}

template <int NF>
__device__ __forceinline__ void pair_wall32(float d0, float d1, float dzf, float dzi, float e2, float h0, float hh,
                                            const float* f, float* acc)
{
    float rho2 = d0 * d0 + d1 * d1;
    float Qf = rho2 + dzf * dzf + e2;
    float rf = rsqrtf(Qf);
    float s3f = rf * rf * rf;
    float h1f = (Qf + e2) * s3f;
    float Qi = rho2 + dzi * dzi + e2;
    float ri = rsqrtf(Qi);
    float ri2 = ri * ri;
    float s3 = ri2 * ri;
    float s5 = s3 * ri2;
    float e2s5 = e2 * s5;
    float A1 = ri + e2 * s3;
    float P1 = -2.0f * hh * (s3 - 3.0f * e2s5);
    float P2 = 6.0f * hh * s5;
    float B1 = -2.0f * h0 * (3.0f * e2s5 + s3);
    float B2 = 2.0f * h0 * s3;
    float B3 = -6.0f * h0 * s5;
    float R1 = 6.0f * h0 * e2s5;
    float cxy = h1f - A1 - P1 + dzi * (R1 - B2);
    float cz = h1f - A1 + P1 + 2.0f * dzi * B2;
    float bz = B1 + dzi * dzi * B3;
    float b3z = B3 * dzi;
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        float f0 = f[3 * r], f1 = f[3 * r + 1], f2 = f[3 * r + 2];
        float w = d0 * f0 + d1 * f1;
        float tf = s3f * (w + dzf * f2);
        float gd = dzi * f2 - w;
        float bp = P2 * gd - s3 * (w + dzi * f2);
        float bv = bp + B2 * f2 + b3z * gd + tf;
        acc[3 * r] += cxy * f0 + d0 * bv;
        acc[3 * r + 1] += cxy * f1 + d1 * bv;
        acc[3 * r + 2] += cz * f2 + dzf * tf + dzi * bp + bz * gd - R1 * w;
    }
}

template <int NF>
__device__ __forceinline__ void pair_brinkman32(double dd0, double dd1, double dd2, float e2, double e2d, float alpha,
                                                float inv_alpha, double alpha_d, const double* bl, const float* f,
                                                float* acc)
{
    float d0 = (float)dd0, d1 = (float)dd1, d2 = (float)dd2;
    float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    float Q = r2 + e2;
    float rs = rsqrtf(Q);
    float z = alpha * (Q * rs);
    float H1, h2;
    if (z < 0.5f) {
        // FP64 island: the series and the radial factors in double precision.
        double Qd = dd0 * dd0 + dd1 * dd1 + dd2 * dd2 + e2d;
        double rsd = rsqrt_pos(Qd);
        double f1, f2;
        brinkman_series_lds(alpha_d * (Qd * rsd), bl, f1, f2);
        double h2d = f2 * (rsd * rsd * rsd);
        H1 = (float)(f1 * rsd + e2d * h2d);
        h2 = (float)h2d;
    } else {
        float E = expf(-z);
        float iz = rs * inv_alpha;
        float iz2 = iz * iz;
        float f1 = E * (1.0f + iz + iz2) - iz2;
        float f2 = 3.0f * iz2 - E * (1.0f + 3.0f * iz + 3.0f * iz2);
        h2 = f2 * (rs * rs * rs);
        H1 = f1 * rs + e2 * h2;
    }
#pragma unroll
    for (int r = 0; r < NF; ++r) {
        float fd = f[3 * r] * d0 + f[3 * r + 1] * d1 + f[3 * r + 2] * d2;
        float t = h2 * fd;
        acc[3 * r] += H1 * f[3 * r] + t * d0;
        acc[3 * r + 1] += H1 * f[3 * r + 1] + t * d1;
        acc[3 * r + 2] += H1 * f[3 * r + 2] + t * d2;
    }
}

// ---------------------------------------------------------------- kernels

struct NbArgs {
    const double* tgt;   // (3, M)
    const double* src;   // (3, N)
    const double* eps;   // eps[j * eps_stride]
    const double* f;     // (3, N) strengths (real part)
    const double* fi;    // (3, N) imaginary part when nf == 2
    double* u_re;        // (3, M)  (nsplit == 1)
    double* u_im;        // (3, M) or null
    double* part;        // (nsplit, nf, 3, M) raw chunk sums (nsplit > 1)
    int64_t M, N, eps_stride, chunk;
    double alpha, inv_alpha;
    double w_re, w_im;   // kernel scale / mu
    int nsplit;
};

template <int NF>
__device__ __forceinline__ void store_result(const NbArgs& a, int64_t i, const double* s)
{
    const int64_t M = a.M;
    if (a.nsplit > 1) {
        double* p = a.part + (int64_t)blockIdx.y * NF * 3 * M;
#pragma unroll
        for (int k = 0; k < 3 * NF; ++k) p[k * M + i] = s[k];
        return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (NF == 1) {
            a.u_re[c * M + i] = a.w_re * s[c];
            if (a.u_im) a.u_im[c * M + i] = a.w_im * s[c];
        } else {
            a.u_re[c * M + i] = a.w_re * s[c] - a.w_im * s[3 + c];
            a.u_im[c * M + i] = a.w_re * s[3 + c] + a.w_im * s[c];
        }
    }
}

// LDS record of one source for the FP64 kernels: x, y, z, eps^2, then for the
// wall kernel 2h, 6h, 2 eps^2 (h = z, the source height), then f (3*NF).
template <int KIND, int NF>
struct Src64 {
    static constexpr int FOFF = KIND == FL_WALL_STOKESLET3D ? 7 : 4;  // offset of the strengths
    static constexpr int W = (FOFF + 3 * NF + 1) / 2 * 2;             // doubles, even for 16-byte reads
};

template <int KIND, int NF>
__device__ __forceinline__ void fill_src64(double* rec, double x, double y, double z, double e2)
{
    rec[0] = x;
    rec[1] = y;
    rec[2] = z;
    rec[3] = e2;
    if (KIND == FL_WALL_STOKESLET3D) {
        rec[4] = 2.0 * z;
        rec[5] = 6.0 * z;
        rec[6] = 2.0 * e2;
    }
}

template <int KIND, int NF, int TPL>
__global__ __launch_bounds__(WG) void nbody64(NbArgs a)
{
    constexpr int W = Src64<KIND, NF>::W;
    constexpr int FO = Src64<KIND, NF>::FOFF;
    __shared__ __attribute__((aligned(16))) double lds[TILE * W];
    __shared__ double bl[BRINK_LDS];
    const int tid = threadIdx.x;
    const int64_t M = a.M, N = a.N;
    if (KIND == FL_BRINKMANLET3D) stage_brinkman(bl);   // read after the tile loop's first barrier

    // the Brinkmanlet pairs run in alpha-scaled coordinates (brinkman_radial)
    const double sc = KIND == FL_BRINKMANLET3D ? a.alpha : 1.0;
    double tx[TPL], ty[TPL], tz[TPL], acc[TPL][3 * NF];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        int64_t i = (int64_t)blockIdx.x * WG * TPL + t * WG + tid;
        int64_t ic = i < M ? i : M - 1;
        tx[t] = sc * a.tgt[ic];
        ty[t] = sc * a.tgt[M + ic];
        tz[t] = sc * a.tgt[2 * M + ic];
#pragma unroll
        for (int k = 0; k < 3 * NF; ++k) acc[t][k] = 0.0;
    }

    const int64_t jb = (int64_t)blockIdx.y * a.chunk;
    const int64_t je = min(N, jb + a.chunk);
    for (int64_t j0 = jb; j0 < je; j0 += TILE) {
        const int64_t j = j0 + tid;
        double rec[W];
#pragma unroll
        for (int k = 0; k < W; ++k) rec[k] = 0.0;
        if (j < je) {
            const double e = sc * a.eps[j * a.eps_stride];
            fill_src64<KIND, NF>(rec, sc * a.src[j], sc * a.src[N + j], sc * a.src[2 * N + j], e * e);
#pragma unroll
            for (int k = 0; k < 3 * NF; ++k)
                rec[FO + k] = k < 3 ? a.f[(int64_t)k * N + j] : a.fi[(int64_t)(k - 3) * N + j];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < W; k += 2)
            *reinterpret_cast<double2*>(&lds[tid * W + k]) = make_double2(rec[k], rec[k + 1]);
        __syncthreads();
        const int nt = (int)min((int64_t)TILE, je - j0);
        for (int s = 0; s < nt; ++s) {
            double r[W];
#pragma unroll
            for (int k = 0; k < W; k += 2) {
                double2 v = *reinterpret_cast<const double2*>(&lds[s * W + k]);
                r[k] = v.x;
                r[k + 1] = v.y;
            }
            if constexpr (KIND == FL_BRINKMANLET3D) {
                pairs_brinkman<NF, TPL>(tx, ty, tz, r, bl, r + FO, acc);
            } else {
#pragma unroll
                for (int t = 0; t < TPL; ++t) {
                    if (KIND == FL_STOKESLET3D)
                        pair_stokeslet<NF>(tx[t] - r[0], ty[t] - r[1], tz[t] - r[2], r[3], r + FO, acc[t]);
                    else
                        pair_wall<NF>(tx[t] - r[0], ty[t] - r[1], tz[t], r[2], r[3], r[4], r[5], r[6], r + FO,
                                      acc[t]);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        int64_t i = (int64_t)blockIdx.x * WG * TPL + t * WG + tid;
        if (i < M) store_result<NF>(a, i, acc[t]);
    }
}

// LDS record for the FP32C kernels: FP64 coordinates, FP32 constants and strengths.
template <int NF>
struct Src32 {
    double x, y, z;
    float e2, hh;
    float f[3 * NF];
    float pad[(3 * NF + 2) % 4 == 0 ? 0 : 4 - (3 * NF + 2) % 4];
};

template <int KIND, int NF, int TPL>
__global__ __launch_bounds__(WG) void nbody32c(NbArgs a)
{
    __shared__ Src32<NF> lds[TILE];
    __shared__ double bl[BRINK_LDS];
    const int tid = threadIdx.x;
    const int64_t M = a.M, N = a.N;
    const float alpha = (float)a.alpha, inv_alpha = (float)a.inv_alpha;
    if (KIND == FL_BRINKMANLET3D) stage_brinkman(bl);   // read after the tile loop's first barrier

    double tx[TPL], ty[TPL], tz[TPL], acc[TPL][3 * NF];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        int64_t i = (int64_t)blockIdx.x * WG * TPL + t * WG + tid;
        int64_t ic = i < M ? i : M - 1;
        tx[t] = a.tgt[ic];
        ty[t] = a.tgt[M + ic];
        tz[t] = a.tgt[2 * M + ic];
#pragma unroll
        for (int k = 0; k < 3 * NF; ++k) acc[t][k] = 0.0;
    }
    const int64_t jb = (int64_t)blockIdx.y * a.chunk;
    const int64_t je = min(N, jb + a.chunk);
    for (int64_t j0 = jb; j0 < je; j0 += TILE) {
        const int64_t j = j0 + tid;
        Src32<NF> rec{};
        if (j < je) {
            rec.x = a.src[j];
            rec.y = a.src[N + j];
            rec.z = a.src[2 * N + j];
            double e = a.eps[j * a.eps_stride];
            rec.e2 = (float)(e * e);
            rec.hh = (float)(rec.z * rec.z);
#pragma unroll
            for (int k = 0; k < 3 * NF; ++k)
                rec.f[k] = (float)(k < 3 ? a.f[(int64_t)k * N + j] : a.fi[(int64_t)(k - 3) * N + j]);
        }
        __syncthreads();
        lds[tid] = rec;
        __syncthreads();
        const int nt = (int)min((int64_t)TILE, je - j0);
        float part[TPL][3 * NF];
#pragma unroll
        for (int t = 0; t < TPL; ++t)
#pragma unroll
            for (int k = 0; k < 3 * NF; ++k) part[t][k] = 0.0f;
        for (int s = 0; s < nt; ++s) {
            const Src32<NF> r = lds[s];
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                double dd0 = tx[t] - r.x, dd1 = ty[t] - r.y, dd2 = tz[t] - r.z;
                if (KIND == FL_STOKESLET3D)
                    pair_stokeslet32<NF>((float)dd0, (float)dd1, (float)dd2, r.e2, r.f, part[t]);
                else if (KIND == FL_WALL_STOKESLET3D)
                    pair_wall32<NF>((float)dd0, (float)dd1, (float)dd2, (float)(tz[t] + r.z), r.e2, (float)r.z,
                                    r.hh, r.f, part[t]);
                else
                    pair_brinkman32<NF>(dd0, dd1, dd2, r.e2, (double)r.e2, alpha, inv_alpha, a.alpha, bl, r.f,
                                        part[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < TPL; ++t)
#pragma unroll
            for (int k = 0; k < 3 * NF; ++k) acc[t][k] += (double)part[t][k];
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        int64_t i = (int64_t)blockIdx.x * WG * TPL + t * WG + tid;
        if (i < M) store_result<NF>(a, i, acc[t]);
    }
}

// Sums the chunk partials in chunk order and applies the scale.
template <int NF>
__global__ __launch_bounds__(WG) void nbody_reduce(NbArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t M = a.M;
    if (i >= M) return;
    double s[3 * NF];
#pragma unroll
    for (int k = 0; k < 3 * NF; ++k) s[k] = 0.0;
    for (int y = 0; y < a.nsplit; ++y) {
        const double* p = a.part + (int64_t)y * NF * 3 * M;
#pragma unroll
        for (int k = 0; k < 3 * NF; ++k) s[k] += p[k * M + i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (NF == 1) {
            a.u_re[c * M + i] = a.w_re * s[c];
            if (a.u_im) a.u_im[c * M + i] = a.w_im * s[c];
        } else {
            a.u_re[c * M + i] = a.w_re * s[c] - a.w_im * s[3 + c];
            a.u_im[c * M + i] = a.w_re * s[3 + c] + a.w_im * s[c];
        }
    }
}

// ---------------------------------------------------------------- symmetric sums
// When the targets are the sources and every source has the same blob width,
// the kernels are reciprocal: G(y, x) = G(x, y)^T (the free-space blocks are
// symmetric and even in d; the wall block maps to its transpose, see
// wall_block).  Then u_i = sum_j G(x_i, x_j) f_j is formed by evaluating each
// unordered pair's block once and applying it both ways:
//   u_i += G(x_i, x_j) f_j     (row part, accumulated in the lane of target i)
//   u_j += G(x_i, x_j)^T f_i   (column part, reduced over the targets i)
// The points are cut into tiles of SB = 512 (256 lanes x 2 targets).  Workgroup
// (K, q) takes target tile K against source tiles L in [K + q c, K + (q+1) c),
// L >= K; the diagonal tile L = K is summed as rows only (each ordered pair
// once, self-pair included).  Column sums are reduced in fixed order: within a
// wave through an LDS transpose (batches of 4 sources), then over the 4 waves,
// and written per (K, L) to colp; row sums per (K, q) to rowp.  nbody_sym_reduce
// adds, for each point, its row partials in q order and then its column
// partials in K order, so results are deterministic.
constexpr int SB = 2 * WG;
constexpr int SBATCH = 4;
// Row pitch of the column-reduction buffer: 68 doubles shifts consecutive rows
// by 8 LDS banks, so the stride-4 reads of the 48 reducing lanes spread evenly.
constexpr int SRS = 68;

struct SymArgs {
    const double* x;   // (3, N)
    const double* f;   // (3, N)
    const double* eps; // one blob width, eps[0]
    double* u_re;      // (3, N)
    double* u_im;      // (3, N) or null
    double* rowp;      // (T, Q, 3, SB)
    double* colp;      // (T (T - 1) / 2, 3, SB)
    int64_t N;
    int T, Q, c;
    double alpha, inv_alpha, w_re, w_im;
};

__device__ __forceinline__ int64_t sym_pair_index(int64_t K, int64_t L, int64_t T)
{
    return K * T - K * (K + 1) / 2 + (L - K - 1);
}

// Block of (target, source) applied to the source strength f (row) and, when
// COL, its transpose applied to the target strength g (added to col).
// H1 I + H2 d d^T applied to the source strength f (row) and, when COL, to
// the target strength g (added to col; the block is symmetric).
template <bool COL>
__device__ __forceinline__ void sym_apply(double d0, double d1, double d2, double H1, double H2, const double* f,
                                          const double* g, double* acc, double* col)
{
    const double t = H2 * fma(d0, f[0], fma(d1, f[1], d2 * f[2]));
    acc[0] = fma(H1, f[0], fma(t, d0, acc[0]));
    acc[1] = fma(H1, f[1], fma(t, d1, acc[1]));
    acc[2] = fma(H1, f[2], fma(t, d2, acc[2]));
    if (COL) {
        const double tg = H2 * fma(d0, g[0], fma(d1, g[1], d2 * g[2]));
        col[0] = fma(H1, g[0], fma(tg, d0, col[0]));
        col[1] = fma(H1, g[1], fma(tg, d1, col[1]));
        col[2] = fma(H1, g[2], fma(tg, d2, col[2]));
    }
}

// The Brinkmanlet blocks of the lane's two targets and one source record r.
template <bool COL>
__device__ __forceinline__ void sym_brinkman2(const double* tx, const double* ty, const double* tz, const double* r,
                                              const double (*g)[3], const double* bl, double (*acc)[3], double* col)
{
    constexpr int FO = Src64<FL_BRINKMANLET3D, 1>::FOFF;
    const double e2 = r[3];
    double d[2][3], H1[2], H2[2], z[2], rs[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        d[t][0] = tx[t] - r[0];
        d[t][1] = ty[t] - r[1];
        d[t][2] = tz[t] - r[2];
        const double Q = fma(d[t][0], d[t][0], fma(d[t][1], d[t][1], fma(d[t][2], d[t][2], e2)));
        rs[t] = rsqrt_pos(Q);
        z[t] = brinkman_radial(Q, rs[t], e2, bl, H1[t], H2[t]);
    }
    if (z[0] < 0.5 || z[1] < 0.5) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
            if (z[t] < 0.5) brinkman_radial_near(z[t], rs[t], e2, bl, H1[t], H2[t]);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) sym_apply<COL>(d[t][0], d[t][1], d[t][2], H1[t], H2[t], r + FO, g[t], acc[t], col);
}

template <int KIND, bool COL>
__device__ __forceinline__ void sym_pair(double d0, double d1, double d2, double z, const double* r,
                                         const double* g, double* acc, double* col)
{
    static_assert(KIND != FL_BRINKMANLET3D, "Brinkmanlet pairs go through sym_brinkman2");
    constexpr int FO = Src64<KIND, 1>::FOFF;
    const double f0 = r[FO], f1 = r[FO + 1], f2 = r[FO + 2];
    if (KIND == FL_WALL_STOKESLET3D) {
        const WallBlock B = wall_block(d0, d1, z, r[2], r[3], r[4], r[5], r[6]);
        const double w = fma(d0, f0, d1 * f1);
        const double v = fma(B.b, w, B.g * f2);
        acc[0] = fma(B.C, f0, fma(d0, v, acc[0]));
        acc[1] = fma(B.C, f1, fma(d1, v, acc[1]));
        acc[2] = fma(B.t, w, fma(B.Z, f2, acc[2]));
        if (COL) {
            const double wg = fma(d0, g[0], d1 * g[1]);
            const double vg = fma(B.b, wg, B.t * g[2]);
            col[0] = fma(B.C, g[0], fma(d0, vg, col[0]));
            col[1] = fma(B.C, g[1], fma(d1, vg, col[1]));
            col[2] = fma(B.g, wg, fma(B.Z, g[2], col[2]));
        }
    } else {
        // Stokeslet: H1 I + H2 d d^T (symmetric)
        const double e2 = r[3];
        const double Q = fma(d0, d0, fma(d1, d1, fma(d2, d2, e2)));
        const double rs = rsqrt_pos(Q);
        const double s3 = rs * rs * rs;
        sym_apply<COL>(d0, d1, d2, (Q + e2) * s3, s3, r + FO, g, acc, col);
    }
}

template <int KIND>
__global__ __launch_bounds__(WG) void nbody64_sym(SymArgs a)
{
    constexpr int W = Src64<KIND, 1>::W;
    constexpr int FO = Src64<KIND, 1>::FOFF;
    __shared__ __attribute__((aligned(16))) double lds[WG * W];          // 256 sources at a time
    __shared__ double red[4][SBATCH * 3 * SRS];  // per wave: [source][comp][lane], rows padded to SRS
    __shared__ double colw[4][64 * 3];                                   // per wave: [source][comp]
    __shared__ double bl[BRINK_LDS];
    const int K = blockIdx.x, q = blockIdx.y;
    const int L0 = K + q * a.c;
    if (L0 >= a.T) return;
    const int L1 = min(a.T, L0 + a.c);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (KIND == FL_BRINKMANLET3D) stage_brinkman(bl);   // read after the tile loop's first barrier
    const int64_t N = a.N;
    // the Brinkmanlet pairs run in alpha-scaled coordinates (brinkman_radial)
    const double sc = KIND == FL_BRINKMANLET3D ? a.alpha : 1.0;
    const double e2 = (sc * a.eps[0]) * (sc * a.eps[0]);

    double tx[2], ty[2], tz[2], g[2][3], acc[2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t i = (int64_t)K * SB + t * WG + tid;
        const int64_t ic = i < N ? i : N - 1;
        tx[t] = sc * a.x[ic];
        ty[t] = sc * a.x[N + ic];
        tz[t] = sc * a.x[2 * N + ic];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[t][c] = i < N ? a.f[c * N + i] : 0.0;
            acc[t][c] = 0.0;
        }
    }

    for (int L = L0; L < L1; ++L) {
        const bool diag = L == K;
        for (int half = 0; half < 2; ++half) {
            const int64_t jbase = (int64_t)L * SB + half * WG;
            {
                const int64_t j = jbase + tid;
                const int64_t jc = j < N ? j : N - 1;
                double rec[W];
#pragma unroll
                for (int k = 0; k < W; ++k) rec[k] = 0.0;
                fill_src64<KIND, 1>(rec, sc * a.x[jc], sc * a.x[N + jc], sc * a.x[2 * N + jc], e2);
#pragma unroll
                for (int k = 0; k < 3; ++k) rec[FO + k] = j < N ? a.f[(int64_t)k * N + j] : 0.0;
                __syncthreads();
#pragma unroll
                for (int k = 0; k < W; k += 2)
                    *reinterpret_cast<double2*>(&lds[tid * W + k]) = make_double2(rec[k], rec[k + 1]);
                __syncthreads();
            }
            if (diag) {
                for (int s = 0; s < WG; ++s) {
                    double r[W];
#pragma unroll
                    for (int k = 0; k < W; k += 2) {
                        const double2 v = *reinterpret_cast<const double2*>(&lds[s * W + k]);
                        r[k] = v.x;
                        r[k + 1] = v.y;
                    }
                    if constexpr (KIND == FL_BRINKMANLET3D) {
                        sym_brinkman2<false>(tx, ty, tz, r, g, bl, acc, nullptr);
                    } else {
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            sym_pair<KIND, false>(tx[t] - r[0], ty[t] - r[1], tz[t] - r[2], tz[t], r, g[t], acc[t],
                                                  nullptr);
                    }
                }
                continue;
            }
            for (int sb = 0; sb < WG / 64; ++sb) {
                for (int s4 = 0; s4 < 64 / SBATCH; ++s4) {
#pragma unroll 1
                    for (int b = 0; b < SBATCH; ++b) {
                        const int s = sb * 64 + s4 * SBATCH + b;
                        double r[W];
#pragma unroll
                        for (int k = 0; k < W; k += 2) {
                            const double2 v = *reinterpret_cast<const double2*>(&lds[s * W + k]);
                            r[k] = v.x;
                            r[k + 1] = v.y;
                        }
                        double col[3] = {0.0, 0.0, 0.0};
                        if constexpr (KIND == FL_BRINKMANLET3D) {
                            sym_brinkman2<true>(tx, ty, tz, r, g, bl, acc, col);
                        } else {
#pragma unroll
                            for (int t = 0; t < 2; ++t)
                                sym_pair<KIND, true>(tx[t] - r[0], ty[t] - r[1], tz[t] - r[2], tz[t], r, g[t],
                                                     acc[t], col);
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) red[wave][(b * 3 + c) * SRS + lane] = col[c];
                    }
                    // the wave's own LDS writes are visible to its later reads (in-order LDS)
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    // 12 (source, comp) sums of 64 values: 4 lanes x 16 values each
                    // (lane part q takes entries q, q + 4, ..., q + 60)
                    const int p = min(lane >> 2, SBATCH * 3 - 1);
                    const double* src = &red[wave][p * SRS + (lane & 3)];
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k < 16; ++k) v += src[4 * k];
                    v += __shfl_xor(v, 1);
                    v += __shfl_xor(v, 2);
                    if ((lane & 3) == 0 && lane < SBATCH * 3 * 4)
                        colw[wave][(s4 * SBATCH + p / 3) * 3 + p % 3] = v;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                __syncthreads();
                if (tid < 192) {
                    const int c = tid >> 6, s = tid & 63;
                    const double v = ((colw[0][s * 3 + c] + colw[1][s * 3 + c]) + colw[2][s * 3 + c]) + colw[3][s * 3 + c];
                    a.colp[(sym_pair_index(K, L, a.T) * 3 + c) * SB + half * WG + sb * 64 + s] = v;
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            a.rowp[(((int64_t)K * a.Q + q) * 3 + c) * SB + t * WG + tid] = acc[t][c];
}

// u_i = w * (row partials of tile K in q order + column partials (K', K) in K' order).
__global__ __launch_bounds__(WG) void nbody_sym_reduce(SymArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t N = a.N;
    if (i >= N) return;
    const int64_t K = i / SB, t = i % SB;
    const int nq = (int)((a.T - K + a.c - 1) / a.c);
    double s[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < nq; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += a.rowp[((K * a.Q + q) * 3 + c) * SB + t];
    for (int64_t Kp = 0; Kp < K; ++Kp)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += a.colp[(sym_pair_index(Kp, K, a.T) * 3 + c) * SB + t];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.u_re[c * N + i] = a.w_re * s[c];
        if (a.u_im) a.u_im[c * N + i] = a.w_im * s[c];
    }
}

int env_int(const char* name, int dflt)
{
    const char* v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

// Targets per lane: 1, or 2 when there are enough targets to give every CU
// two register-blocked workgroups anyway.  FL_NBODY_TPL=1|2 overrides (tuning).
int pick_tpl(int64_t M)
{
    const int forced = env_int("FL_NBODY_TPL", 0);
    if (forced == 1 || forced == 2) return forced;
    return M >= (int64_t)WG * 2 * 2 * device_cus() ? 2 : 1;
}

// Number of source chunks: enough workgroups to fill the device (>= 4 per CU)
// while every chunk keeps at least 2048 sources.
int pick_split(int64_t M, int64_t N, int tpl)
{
    int64_t wgx = (M + (int64_t)WG * tpl - 1) / ((int64_t)WG * tpl);
    int64_t want = 4LL * device_cus();
    int64_t s = (want + wgx - 1) / wgx;
    int64_t smax = std::max<int64_t>(1, N / 2048);
    s = std::min<int64_t>(s, smax);
    s = std::min<int64_t>(s, 256);
    return (int)std::max<int64_t>(1, s);
}

__global__ __launch_bounds__(WG) void zero_kernel(double* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Symmetric plan: T tiles of SB points, c tiles per workgroup along a row,
// Q = ceil(T / c) chunks (grid.y).  Used only when there are enough tile pairs
// to give every CU several workgroups (T (T + 1) / 2 >= 4 CUs), when its
// workspace (one column partial per tile pair, about 0.023 N^2 bytes) stays
// within SYM_MAX_BYTES (N up to about 428,000; larger sums take the general
// kernel, whose workspace is O(N)), and not switched off by
// FL_EVAL_SYMMETRIC=0 (measurement switch).
constexpr size_t SYM_MAX_BYTES = size_t(4) << 30;
struct SymPlan {
    int T = 0, c = 0, Q = 0;
    size_t bytes = 0;
    bool ok = false;
};

SymPlan sym_plan(int64_t N)
{
    SymPlan p;
    if (env_int("FL_EVAL_SYMMETRIC", 1) == 0 || N <= 0) return p;
    const int64_t T = (N + SB - 1) / SB;
    const int64_t pairs = T * (T + 1) / 2;
    const int64_t cus = device_cus();
    if (pairs < 4 * cus || T > 65535) return p;
    const int64_t c = std::min<int64_t>(16, std::max<int64_t>(1, pairs / (16 * cus)));
    p.T = (int)T;
    p.c = (int)c;
    p.Q = (int)((T + c - 1) / c);
    const size_t rowb = align_up((size_t)T * p.Q * 3 * SB * sizeof(double), 256);
    const size_t colb = (size_t)(T * (T - 1) / 2) * 3 * SB * sizeof(double);
    p.bytes = rowb + colb;
    p.ok = p.bytes <= SYM_MAX_BYTES;
    return p;
}

template <int KIND, int NF, int TPL>
void launch_main(const NbArgs& a, dim3 grid, int precision, hipStream_t stream)
{
    if (precision == FL_FP64)
        hipLaunchKernelGGL((nbody64<KIND, NF, TPL>), grid, dim3(WG), 0, stream, a);
    else
        hipLaunchKernelGGL((nbody32c<KIND, NF, TPL>), grid, dim3(WG), 0, stream, a);
}

template <int NF, int TPL>
void launch_kind(int kind, const NbArgs& a, dim3 grid, int precision, hipStream_t stream)
{
    if (kind == FL_STOKESLET3D) launch_main<FL_STOKESLET3D, NF, TPL>(a, grid, precision, stream);
    else if (kind == FL_WALL_STOKESLET3D) launch_main<FL_WALL_STOKESLET3D, NF, TPL>(a, grid, precision, stream);
    else launch_main<FL_BRINKMANLET3D, NF, TPL>(a, grid, precision, stream);
}


// ---------------------------------------------------------------- barycentric Lagrange treecode
// Fast summation of the same kernel sums (a particle-cluster treecode with
// barycentric Lagrange interpolation: Wang, Tlupova & Krasny, SIAM J. Sci.
// Comput. 42 (2020) A2993; kernel-independent, so the regularized kernels and
// the wall's image system are interpolated as they are).  For a cluster C of
// sources with box [lo, hi]^3 and Chebyshev points of the second kind
// s_k = (lo + hi)/2 + (hi - lo)/2 cos(k pi / n), k = 0 .. n, in each
// dimension, the (n + 1)^3 proxy sources at s_(k1,k2,k3) carry the strengths
//   F_k = sum_{y_j in C} L_k1(y_j1) L_k2(y_j2) L_k3(y_j3) f_j,
//   L_k(x) = (w_k / (x - s_k)) / sum_l (w_l / (x - s_l)),   w_k = (-1)^k (1/2 at k = 0, n),
// (L_k(x) = 1 when x is s_k), and a target far from C in the tree's
// acceptance criterion sums the kernel over the proxies instead of over C.
// The host builds the tree and each target block's list of segments (runs of
// real sources or of one cluster's proxies); the kernels below compute the
// proxy strengths and then walk the lists with the direct sum's pair code.

constexpr int BLTC_NMAX = 12;   // largest interpolation degree

constexpr int BLTC_TS = 64;    // sources staged per tile of the moments kernel

// LDS of the moments kernel at n + 1 = N1 points per dimension: a tile of
// source records (f in 6 slots, then L_k(y_1), L_k(y_2), L_k(y_3), N1 each;
// an even count, for 16-byte reads), reused for the pairwise group sums.
template <int N1>
constexpr int bltc_rw() { return (6 + 3 * N1 + 1) / 2 * 2; }
template <int NF, int N1>
constexpr size_t bltc_moments_lds()
{
    constexpr size_t tile = (size_t)BLTC_TS * bltc_rw<N1>();
    constexpr size_t sums = (size_t)(WG / (N1 * N1) / 2) * N1 * N1 * N1 * 3 * NF;
    return (tile > sums ? tile : sums) * sizeof(double);
}

struct BltcWeightArgs {
    const int64_t* cl;     // per cluster: source begin, source end, proxy base, first chunk, chunk count
    const int64_t* chunk;  // per chunk: cluster, source begin, source end
    const double* box;     // per cluster: lo x, y, z, hi x, y, z
    const double* src;     // (3, N) sorted sources
    const double* f;       // (3, N) strengths (real part)
    const double* fi;      // (3, N) imaginary part or null
    double* px;            // (3, P) proxy coordinates
    double* pf;            // (3, P) proxy strengths
    double* pfi;           // (3, P) imaginary part or null
    double* ws;            // per chunk: (3 NF, (n + 1)^3) partial strengths
    int64_t N, P;
    int degree;
};

// Chebyshev point k (of the second kind, degree n) of dimension d of a box.
__device__ __forceinline__ double bltc_node(const double* bx, int d, int k, int n)
{
    return 0.5 * (bx[d] + bx[3 + d]) + 0.5 * (bx[3 + d] - bx[d]) * cos(M_PI * k / n);
}

// Partial proxy strengths of one chunk of a cluster's sources (one workgroup
// per chunk, so a cluster of half a million sources is summed by a hundred
// workgroups rather than one).  The strengths are a product
//   F[k1][k2 k3][c] = sum_j L_k1(y_j1) (L_k2(y_j2) L_k3(y_j3) f_jc),
// so each lane owns one (k2, k3) and all n + 1 values of k1: per source it
// forms the bracket once (1 + 3 NF multiplies) and adds (n + 1) 3 NF FMAs,
// the L_k1 and f values being LDS broadcasts.  The (n + 1)^2 lanes of a
// (k2, k3) set make a group; the WG / (n + 1)^2 groups take the tile's
// sources in turn and are summed pairwise through LDS in a fixed tree, so
// the result does not depend on timing.  Sources are staged 64 at a time:
// 192 lanes compute the basis values of one dimension of one source (the
// barycentric formula, exact 1 / 0 where the coordinate is a node), the last
// 64 load the strengths.
template <int NF, int N1>
__global__ __launch_bounds__(WG) void bltc_moments_kernel(BltcWeightArgs a)
{
    constexpr int TS = BLTC_TS, RW = bltc_rw<N1>(), NC = 3 * NF, n = N1 - 1;
    constexpr int G = N1 * N1, NG = WG / G;
    static_assert(4 * TS == WG, "staging: three dimensions and the strengths, one source per lane each");
    static_assert(NG >= 1, "a group of (n + 1)^2 lanes fits a workgroup");
    extern __shared__ __attribute__((aligned(16))) double lds[];   // tile records; then the group sums
    __shared__ double nodes[3 * N1], bw[N1];
    const int tid = threadIdx.x;
    const int64_t ch = blockIdx.x;
    const int64_t c = a.chunk[3 * ch], jb = a.chunk[3 * ch + 1], je = a.chunk[3 * ch + 2];
    const double* bx = a.box + 6 * c;
    if (tid < 3 * N1) nodes[tid] = bltc_node(bx, tid / N1, tid % N1, n);
    if (tid < N1) bw[tid] = ((tid & 1) ? -1.0 : 1.0) * ((tid == 0 || tid == n) ? 0.5 : 1.0);
    const int g = tid / G, p = tid % G, k2 = p / N1, k3 = p % N1;
    const bool act = g < NG;
    double acc[N1][NC];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
        for (int k = 0; k < NC; ++k) acc[k1][k] = 0.0;
    for (int64_t j0 = jb; j0 < je; j0 += TS) {
        const int cnt = (int)min((int64_t)TS, je - j0);
        __syncthreads();   // the previous tile is consumed (and the nodes are staged)
        if (tid < 3 * TS) {
            const int d = tid / TS, i = tid % TS;
            if (i < cnt) {
                double* L = lds + i * RW + 6 + d * N1;
                const double x = a.src[d * a.N + j0 + i];
                double t[N1], sum = 0.0;
                int hit = -1;
#pragma unroll
                for (int k = 0; k < N1; ++k) {
                    const double dx = x - nodes[d * N1 + k];
                    if (dx == 0.0 && hit < 0) hit = k;
                    t[k] = dx != 0.0 ? bw[k] / dx : 0.0;
                    sum += t[k];
                }
                const double inv = hit < 0 ? 1.0 / sum : 0.0;
#pragma unroll
                for (int k = 0; k < N1; ++k) L[k] = hit < 0 ? t[k] * inv : (k == hit ? 1.0 : 0.0);
            }
        } else {
            const int i = tid - 3 * TS;
            if (i < cnt) {
                double* r = lds + i * RW;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    r[k] = a.f[k * a.N + j0 + i];
                    if (NF == 2) r[3 + k] = a.fi[k * a.N + j0 + i];
                }
            }
        }
        __syncthreads();
        if (act) {
            for (int i = g; i < cnt; i += NG) {
                const double* r = lds + i * RW;
                const double w = r[6 + N1 + k2] * r[6 + 2 * N1 + k3];
                double fw[NC];
#pragma unroll
                for (int k = 0; k < NC; ++k) fw[k] = w * r[k];
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    const double l = r[6 + k1];
#pragma unroll
                    for (int k = 0; k < NC; ++k) acc[k1][k] = fma(l, fw[k], acc[k1][k]);
                }
            }
        }
    }
    // groups g and g + s (g a multiple of 2 s) are summed into g, s = 1, 2, 4 ...
    constexpr int SLOT = G * N1 * NC;
#pragma unroll
    for (int s = 1; s < NG; s *= 2) {
        __syncthreads();
        if (act && g % (2 * s) == s) {
            double* o = lds + (g / (2 * s)) * SLOT + p;
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
                for (int k = 0; k < NC; ++k) o[(k1 * NC + k) * G] = acc[k1][k];
        }
        __syncthreads();
        if (act && g % (2 * s) == 0 && g + s < NG) {
            const double* o = lds + (g / (2 * s)) * SLOT + p;
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
                for (int k = 0; k < NC; ++k) acc[k1][k] += o[(k1 * NC + k) * G];
        }
    }
    if (g == 0) {
        constexpr int Q = G * N1;
        double* o = a.ws + ch * NC * Q + p;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
            for (int k = 0; k < NC; ++k) o[k * Q + k1 * G] = acc[k1][k];
    }
}

// One workgroup per cluster: the proxies' coordinates (the Chebyshev points
// of its box, proxy q = (k1 (n + 1) + k2) (n + 1) + k3) and strengths, the
// sum of its chunks' partial strengths in chunk order.
template <int NF>
__global__ __launch_bounds__(WG) void bltc_proxy_kernel(BltcWeightArgs a)
{
    constexpr int NC = 3 * NF;
    const int64_t c = blockIdx.x;
    const int n = a.degree, n1 = n + 1, Q = n1 * n1 * n1;
    const int64_t pb = a.cl[5 * c + 2], c0 = a.cl[5 * c + 3], nch = a.cl[5 * c + 4];
    const double* bx = a.box + 6 * c;
    for (int q = threadIdx.x; q < Q; q += WG) {
        const int kk[3] = {q / (n1 * n1), (q / n1) % n1, q % n1};
        double s[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) s[k] = 0.0;
        for (int64_t h = c0; h < c0 + nch; ++h) {
            const double* w = a.ws + h * NC * Q + q;
#pragma unroll
            for (int k = 0; k < NC; ++k) s[k] += w[k * Q];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a.px[d * a.P + pb + q] = bltc_node(bx, d, kk[d], n);
            a.pf[d * a.P + pb + q] = s[d];
            if (NF == 2) a.pfi[d * a.P + pb + q] = s[3 + d];
        }
    }
}

struct BltcEvalArgs {
    const double* tgt;     // (3, M) sorted targets
    const int64_t* tidx;   // original index of each sorted target
    const int64_t* blk;    // per target block: target begin, count (<= WG BLTC_TPL = 256), segment begin, segment end
    const int64_t* seg;    // per segment: kind (0 sources, 1 proxies), begin, count
    const double* src;     // (3, N) sorted sources, strengths f / fi
    const double* f;
    const double* fi;
    const double* px;      // (3, P) proxies, strengths pf / pfi
    const double* pf;
    const double* pfi;
    double* u_re;          // (3, M) in the original target order
    double* u_im;
    int64_t M, N, P;
    double e2;             // (scaled) eps^2, one blob width
    double sc;             // coordinate scale (alpha for the Brinkmanlet)
    double w_re, w_im;
};

// One workgroup per target block (BLTC_TPL targets per lane, lane-strided as
// in nbody64): its segments in list order, each staged through LDS a tile at
// a time and summed with the direct kernel's pair code, so a list that holds
// every source directly gives the direct sum's terms in another order.  One
// target per lane: with two, the blocks (and workgroups) halve and the
// uneven lists balance worse (90 against 79 ms at 262,144 points, degree 8).
constexpr int BLTC_TPL = 1;

template <int KIND, int NF>
__global__ __launch_bounds__(WG) void bltc_eval_kernel(BltcEvalArgs a)
{
    constexpr int W = Src64<KIND, NF>::W;
    constexpr int FO = Src64<KIND, NF>::FOFF;
    constexpr int TPL = BLTC_TPL;
    __shared__ __attribute__((aligned(16))) double lds[TILE * W];
    __shared__ double bl[BRINK_LDS];
    const int tid = threadIdx.x;
    if (KIND == FL_BRINKMANLET3D) stage_brinkman(bl);
    const int64_t* bk = a.blk + 4 * (int64_t)blockIdx.x;
    const int64_t t0 = bk[0], tc = bk[1], s0 = bk[2], s1 = bk[3];
    double tx[TPL], ty[TPL], tz[TPL], acc[TPL][3 * NF];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const int64_t i = t0 + min((int64_t)(t * WG + tid), tc - 1);
        tx[t] = a.sc * a.tgt[i];
        ty[t] = a.sc * a.tgt[a.M + i];
        tz[t] = a.sc * a.tgt[2 * a.M + i];
#pragma unroll
        for (int k = 0; k < 3 * NF; ++k) acc[t][k] = 0.0;
    }
    for (int64_t sg = s0; sg < s1; ++sg) {
        const bool prox = a.seg[3 * sg] != 0;
        const int64_t jb = a.seg[3 * sg + 1], cnt = a.seg[3 * sg + 2];
        const double* X = prox ? a.px : a.src;
        const double* F = prox ? a.pf : a.f;
        const double* FI = prox ? a.pfi : a.fi;
        const int64_t ld = prox ? a.P : a.N;
        for (int64_t j0 = 0; j0 < cnt; j0 += TILE) {
            double rec[W];
#pragma unroll
            for (int k = 0; k < W; ++k) rec[k] = 0.0;
            if (j0 + tid < cnt) {
                const int64_t j = jb + j0 + tid;
                fill_src64<KIND, NF>(rec, a.sc * X[j], a.sc * X[ld + j], a.sc * X[2 * ld + j], a.e2);
#pragma unroll
                for (int k = 0; k < 3 * NF; ++k) rec[FO + k] = k < 3 ? F[(int64_t)k * ld + j] : FI[(int64_t)(k - 3) * ld + j];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < W; k += 2)
                *reinterpret_cast<double2*>(&lds[tid * W + k]) = make_double2(rec[k], rec[k + 1]);
            __syncthreads();
            const int nt = (int)min((int64_t)TILE, cnt - j0);
            for (int s = 0; s < nt; ++s) {
                double r[W];
#pragma unroll
                for (int k = 0; k < W; k += 2) {
                    double2 v = *reinterpret_cast<const double2*>(&lds[s * W + k]);
                    r[k] = v.x;
                    r[k + 1] = v.y;
                }
                if constexpr (KIND == FL_BRINKMANLET3D) {
                    pairs_brinkman<NF, TPL>(tx, ty, tz, r, bl, r + FO, acc);
                } else {
#pragma unroll
                    for (int t = 0; t < TPL; ++t) {
                        if (KIND == FL_STOKESLET3D)
                            pair_stokeslet<NF>(tx[t] - r[0], ty[t] - r[1], tz[t] - r[2], r[3], r + FO, acc[t]);
                        else
                            pair_wall<NF>(tx[t] - r[0], ty[t] - r[1], tz[t], r[2], r[3], r[4], r[5], r[6], r + FO,
                                          acc[t]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        if (t * WG + tid < tc) {
            const int64_t o = a.tidx[t0 + t * WG + tid], M = a.M;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (NF == 1) {
                    a.u_re[c * M + o] = a.w_re * acc[t][c];
                    if (a.u_im) a.u_im[c * M + o] = a.w_im * acc[t][c];
                } else {
                    a.u_re[c * M + o] = a.w_re * acc[t][c] - a.w_im * acc[t][3 + c];
                    a.u_im[c * M + o] = a.w_re * acc[t][3 + c] + a.w_im * acc[t][c];
                }
            }
        }
    }
}

}  // namespace

double kernel_scale(int kind)
{
    const double pi = 3.141592653589793;
    return kind == FL_BRINKMANLET3D ? 1.0 / (4.0 * pi) : 1.0 / (8.0 * pi);
}

size_t eval_workspace_size(int64_t M, int64_t N, int nf, int precision)
{
    (void)precision;
    if (M <= 0 || N <= 0) return 0;
    int tpl = pick_tpl(M);
    int s = pick_split(M, N, tpl);
    if (s <= 1) return 0;
    return (size_t)s * nf * 3 * (size_t)M * sizeof(double);
}

size_t eval_sym_workspace_size(int64_t N, int nf, int precision)
{
    if (nf != 1 || precision != FL_FP64) return 0;
    const SymPlan p = sym_plan(N);
    return p.ok ? p.bytes : 0;
}

int eval_launch(int kind, double alpha, int64_t M, int64_t N, const double* tgt, const double* src,
                const double* eps, int64_t eps_stride, const double* f, const double* fi, double w_re, double w_im,
                double* u_re, double* u_im, int precision, void* ws, size_t ws_bytes, hipStream_t stream)
{
    const int nf = fi ? 2 : 1;
    FL_CHECK(kind >= 0 && kind <= 2, "unknown kernel kind %d", kind);
    FL_CHECK(precision == FL_FP64 || precision == FL_FP32C, "unknown precision %d", precision);
    FL_CHECK(M >= 0 && N >= 0, "negative size");
    FL_CHECK(eps_stride == 0 || eps_stride == 1, "eps_stride must be 0 or 1");
    FL_CHECK(!(nf == 2 || w_im != 0.0) || u_im != nullptr, "complex result needs u_im");
    if (M == 0) return FL_OK;
    if (N == 0) {
        // a fill kernel rather than hipMemsetAsync, so the call can be captured in a graph
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((3 * M + WG - 1) / WG)), dim3(WG), 0, stream, u_re, 3 * M);
        FL_LAUNCHED();
        if (u_im) {
            hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((3 * M + WG - 1) / WG)), dim3(WG), 0, stream, u_im, 3 * M);
            FL_LAUNCHED();
        }
        return FL_OK;
    }
    FL_CHECK(tgt && src && eps && f && u_re, "null pointer argument");
    FL_CHECK(kind != FL_BRINKMANLET3D || (alpha >= 0.0 && std::isfinite(alpha)), "alpha must be finite and >= 0");
    if (kind == FL_BRINKMANLET3D && precision == FL_FP64) {
        if (alpha < 1e-150) {
            // the Brinkmanlet at alpha = 0 is the Stokeslet: b1 = b2 = 1/2 at z = 0,
            // so its bracket is half the Stokeslet's (exact in the scale); below
            // 1e-150 (where scaled coordinates would underflow) it differs from
            // it by O(alpha r), far below rounding
            kind = FL_STOKESLET3D;
            w_re *= 0.5;
            w_im *= 0.5;
        } else {
            // the FP64 Brinkmanlet kernels sum in alpha-scaled coordinates:
            // the sums come out divided by alpha
            w_re *= alpha;
            w_im *= alpha;
        }
    }
    if (tgt == src && M == N && eps_stride == 0 && nf == 1 && precision == FL_FP64) {
        const SymPlan p = sym_plan(N);
        if (p.ok && ws && ws_bytes >= p.bytes) {
            SymArgs s;
            s.x = src;
            s.f = f;
            s.eps = eps;
            s.u_re = u_re;
            s.u_im = u_im;
            s.rowp = static_cast<double*>(ws);
            s.colp = s.rowp + align_up((size_t)p.T * p.Q * 3 * SB * sizeof(double), 256) / sizeof(double);
            s.N = N;
            s.T = p.T;
            s.Q = p.Q;
            s.c = p.c;
            s.alpha = alpha;
            s.inv_alpha = alpha > 0.0 ? 1.0 / alpha : 0.0;
            s.w_re = w_re;
            s.w_im = w_im;
            dim3 grid((unsigned)p.T, (unsigned)p.Q);
            if (kind == FL_STOKESLET3D) hipLaunchKernelGGL(nbody64_sym<FL_STOKESLET3D>, grid, dim3(WG), 0, stream, s);
            else if (kind == FL_WALL_STOKESLET3D)
                hipLaunchKernelGGL(nbody64_sym<FL_WALL_STOKESLET3D>, grid, dim3(WG), 0, stream, s);
            else hipLaunchKernelGGL(nbody64_sym<FL_BRINKMANLET3D>, grid, dim3(WG), 0, stream, s);
            FL_LAUNCHED();
            hipLaunchKernelGGL(nbody_sym_reduce, dim3((unsigned)((N + WG - 1) / WG)), dim3(WG), 0, stream, s);
            FL_LAUNCHED();
            return FL_OK;
        }
    }
    const int tpl = pick_tpl(M);
    const int nsplit = pick_split(M, N, tpl);
    const size_t need = eval_workspace_size(M, N, nf, precision);
    if (ws_bytes < need || (need > 0 && ws == nullptr))
        return set_error(FL_EWORKSPACE, "fl_eval needs %zu workspace bytes, got %zu", need, ws_bytes);

    NbArgs a;
    a.tgt = tgt;
    a.src = src;
    a.eps = eps;
    a.f = f;
    a.fi = fi;
    a.u_re = u_re;
    a.u_im = u_im;
    a.part = static_cast<double*>(ws);
    a.M = M;
    a.N = N;
    a.eps_stride = eps_stride;
    a.chunk = (N + nsplit - 1) / nsplit;
    a.chunk = (a.chunk + TILE - 1) / TILE * TILE;
    a.nsplit = (int)((N + a.chunk - 1) / a.chunk);
    a.alpha = alpha;
    a.inv_alpha = alpha > 0.0 ? 1.0 / alpha : 0.0;
    a.w_re = w_re;
    a.w_im = w_im;
    dim3 grid((unsigned)((M + (int64_t)WG * tpl - 1) / ((int64_t)WG * tpl)), (unsigned)a.nsplit);
    if (nf == 1) {
        if (tpl == 2) launch_kind<1, 2>(kind, a, grid, precision, stream);
        else launch_kind<1, 1>(kind, a, grid, precision, stream);
    } else {
        if (tpl == 2) launch_kind<2, 2>(kind, a, grid, precision, stream);
        else launch_kind<2, 1>(kind, a, grid, precision, stream);
    }
    FL_LAUNCHED();
    if (a.nsplit > 1) {
        dim3 rg((unsigned)((M + WG - 1) / WG));
        if (nf == 1) hipLaunchKernelGGL(nbody_reduce<1>, rg, dim3(WG), 0, stream, a);
        else hipLaunchKernelGGL(nbody_reduce<2>, rg, dim3(WG), 0, stream, a);
        FL_LAUNCHED();
    }
    return FL_OK;
}


template <int N1>
void bltc_moments_launch(int nf, int64_t nchunk, const BltcWeightArgs& a, hipStream_t stream)
{
    constexpr size_t lds1 = bltc_moments_lds<1, N1>(), lds2 = bltc_moments_lds<2, N1>();
    static_assert(lds2 <= 65536, "moments kernel LDS");
    if (nf == 2) hipLaunchKernelGGL((bltc_moments_kernel<2, N1>), dim3((unsigned)nchunk), dim3(WG), lds2, stream, a);
    else hipLaunchKernelGGL((bltc_moments_kernel<1, N1>), dim3((unsigned)nchunk), dim3(WG), lds1, stream, a);
}

// Proxy strengths of ncl clusters at interpolation degree n (1 .. BLTC_NMAX):
// cl per cluster (source begin, end, proxy base, first chunk, chunk count),
// box (lo xyz, hi xyz); chunk per chunk (cluster, source begin, end), the
// chunks of a cluster consecutive; ws holds nchunk 3 nf (n + 1)^3 doubles.
int bltc_weights_launch(int64_t ncl, const int64_t* cl, const double* box, int degree, int64_t nchunk,
                        const int64_t* chunk, int64_t N, const double* src, const double* f, const double* fi,
                        int64_t P, double* px, double* pf, double* pfi, double* ws, hipStream_t stream)
{
    FL_CHECK(degree >= 1 && degree <= BLTC_NMAX, "treecode degree must be 1 .. %d", BLTC_NMAX);
    FL_CHECK(ncl >= 0 && nchunk >= 0 && N >= 0 && P >= 0, "negative size");
    FL_CHECK(!fi == !pfi, "imaginary strengths need imaginary proxy strengths");
    if (ncl == 0) return FL_OK;
    FL_CHECK(cl && box && px && pf, "null pointer argument");
    FL_CHECK(nchunk == 0 || (chunk && src && f && ws), "null pointer argument");
    BltcWeightArgs a{cl, chunk, box, src, f, fi, px, pf, pfi, ws, N, P, degree};
    const int nf = fi ? 2 : 1;
    if (nchunk > 0) {
        switch (degree) {
        case 1: bltc_moments_launch<2>(nf, nchunk, a, stream); break;
        case 2: bltc_moments_launch<3>(nf, nchunk, a, stream); break;
        case 3: bltc_moments_launch<4>(nf, nchunk, a, stream); break;
        case 4: bltc_moments_launch<5>(nf, nchunk, a, stream); break;
        case 5: bltc_moments_launch<6>(nf, nchunk, a, stream); break;
        case 6: bltc_moments_launch<7>(nf, nchunk, a, stream); break;
        case 7: bltc_moments_launch<8>(nf, nchunk, a, stream); break;
        case 8: bltc_moments_launch<9>(nf, nchunk, a, stream); break;
        case 9: bltc_moments_launch<10>(nf, nchunk, a, stream); break;
        case 10: bltc_moments_launch<11>(nf, nchunk, a, stream); break;
        case 11: bltc_moments_launch<12>(nf, nchunk, a, stream); break;
        default: bltc_moments_launch<13>(nf, nchunk, a, stream); break;
        }
        FL_LAUNCHED();
    }
    if (nf == 2) hipLaunchKernelGGL(bltc_proxy_kernel<2>, dim3((unsigned)ncl), dim3(WG), 0, stream, a);
    else hipLaunchKernelGGL(bltc_proxy_kernel<1>, dim3((unsigned)ncl), dim3(WG), 0, stream, a);
    FL_LAUNCHED();
    return FL_OK;
}

// The treecode's kernel sums over the target blocks' segment lists.
int bltc_eval_launch(int kind, double alpha, int64_t M, const double* tgt, const int64_t* tidx, int64_t nblk,
                     const int64_t* blk, const int64_t* seg, int64_t N, const double* src, const double* f,
                     const double* fi, int64_t P, const double* px, const double* pf, const double* pfi, double eps,
                     double w_re, double w_im, double* u_re, double* u_im, hipStream_t stream)
{
    const int nf = fi ? 2 : 1;
    FL_CHECK(kind >= 0 && kind <= 2, "unknown kernel kind %d", kind);
    FL_CHECK(M >= 0 && N >= 0 && P >= 0 && nblk >= 0, "negative size");
    FL_CHECK(!(nf == 2 || w_im != 0.0) || u_im != nullptr, "complex result needs u_im");
    FL_CHECK(!fi == !pfi, "imaginary strengths need imaginary proxy strengths");
    FL_CHECK(eps > 0.0 && std::isfinite(eps), "eps must be finite and > 0");
    if (M == 0 || nblk == 0) return FL_OK;
    FL_CHECK(tgt && tidx && blk && seg && u_re, "null pointer argument");
    FL_CHECK(kind != FL_BRINKMANLET3D || (alpha >= 0.0 && std::isfinite(alpha)), "alpha must be finite and >= 0");
    double sc = 1.0;
    if (kind == FL_BRINKMANLET3D) {
        if (alpha < 1e-150) {   // the Stokeslet, as in eval_launch
            kind = FL_STOKESLET3D;
            w_re *= 0.5;
            w_im *= 0.5;
        } else {
            sc = alpha;
            w_re *= alpha;
            w_im *= alpha;
        }
    }
    BltcEvalArgs a{tgt, tidx, blk, seg, src, f, fi, px, pf, pfi, u_re, u_im, M, N, P, sc * sc * eps * eps, sc, w_re, w_im};
    const dim3 grid((unsigned)nblk);
    if (nf == 1) {
        if (kind == FL_STOKESLET3D) hipLaunchKernelGGL((bltc_eval_kernel<FL_STOKESLET3D, 1>), grid, dim3(WG), 0, stream, a);
        else if (kind == FL_WALL_STOKESLET3D) hipLaunchKernelGGL((bltc_eval_kernel<FL_WALL_STOKESLET3D, 1>), grid, dim3(WG), 0, stream, a);
        else hipLaunchKernelGGL((bltc_eval_kernel<FL_BRINKMANLET3D, 1>), grid, dim3(WG), 0, stream, a);
    } else {
        if (kind == FL_STOKESLET3D) hipLaunchKernelGGL((bltc_eval_kernel<FL_STOKESLET3D, 2>), grid, dim3(WG), 0, stream, a);
        else if (kind == FL_WALL_STOKESLET3D) hipLaunchKernelGGL((bltc_eval_kernel<FL_WALL_STOKESLET3D, 2>), grid, dim3(WG), 0, stream, a);
        else hipLaunchKernelGGL((bltc_eval_kernel<FL_BRINKMANLET3D, 2>), grid, dim3(WG), 0, stream, a);
    }
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace fl
