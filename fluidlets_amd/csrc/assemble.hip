// Dense kernel matrices: K[3i+a, 3j+b] = G_ab(x_i, y_j; eps_j) / mu.
//
// HBM-write bound (72 N^2 bytes in FP64, +36 N^2 for the FP32 copy written in
// the same pass).  A workgroup owns 256 consecutive sources (columns 3j..3j+2)
// and RB consecutive targets; each lane forms the 3x3 block of its source for
// every target and writes three contiguous values per row, so a wave writes
// 64 * 3 consecutive entries of each row.  Per-pair algebra follows
// kernels/stokeslet3d.py:32-50, kernels/images.py:26-89 (full image matrix,
// element pattern of images.py:48-74) and kernels/brinkmanlet3d.py:74-104.
#include "fl_internal.h"

namespace fl {
namespace {

constexpr int WG = 256;
constexpr int RB = 4;  // targets per workgroup
constexpr int NTERMS = 24;

struct SeriesCoef {
    double c1[NTERMS];
    double c2[NTERMS];
};

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

__constant__ SeriesCoef kSeriesA = make_series();

// Full 3x3 block G (row-major) in units of the kernel scale (1/(8 pi) for the
// Stokeslet families, 1/(4 pi) for the Brinkmanlet).
template <int KIND>
__device__ __forceinline__ void block3x3(double tx, double ty, double tz, double sx, double sy, double sz, double e2,
                                         double alpha, double inv_alpha, double* G)
{
    double d0 = tx - sx, d1 = ty - sy, d2 = tz - sz;
    if (KIND == FL_STOKESLET3D || KIND == FL_WALL_STOKESLET3D) {
        double Q = d0 * d0 + d1 * d1 + d2 * d2 + e2;
        double rs = rsqrt(Q);
        double s3 = rs * rs * rs;
        double h1 = (Q + e2) * s3;
        G[0] = h1 + s3 * d0 * d0;
        G[1] = s3 * d0 * d1;
        G[2] = s3 * d0 * d2;
        G[3] = G[1];
        G[4] = h1 + s3 * d1 * d1;
        G[5] = s3 * d1 * d2;
        G[6] = G[2];
        G[7] = G[5];
        G[8] = h1 + s3 * d2 * d2;
        if (KIND == FL_WALL_STOKESLET3D) {
            // image system, coefficients doubled to share the 1/(8 pi) scale
            double h0 = sz, hh = sz * sz;
            double e0 = d0, e1 = d1, e3 = tz + h0;
            double Qi = e0 * e0 + e1 * e1 + e3 * e3 + e2;
            double ri = rsqrt(Qi);
            double ri2 = ri * ri;
            double q3 = ri2 * ri;
            double q5 = q3 * ri2;
            double e2q5 = e2 * q5;
            double A1 = ri + e2 * q3;
            double A2 = q3;
            double P1 = -2.0 * hh * (q3 - 3.0 * e2q5);
            double P2 = 6.0 * hh * q5;
            double B1 = -2.0 * h0 * (3.0 * e2q5 + q3);
            double B2 = 2.0 * h0 * q3;
            double B3 = -6.0 * h0 * q5;
            double R1 = 6.0 * h0 * e2q5;
            double dd[3] = {e0, e1, e3};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double delta = i == j ? 1.0 : 0.0;
                    double a = -(delta * A1 + dd[i] * dd[j] * A2);
                    double sg = j == 2 ? 1.0 : -1.0;
                    a += sg * (delta * P1 + dd[i] * dd[j] * P2);
                    G[3 * i + j] += a;
                }
            G[0] += -e3 * (B2 + e0 * e0 * B3) + e3 * R1;
            G[1] += -e0 * e1 * e3 * B3;
            G[2] += e0 * (B2 + e3 * e3 * B3);
            G[3] += -e1 * e0 * e3 * B3;
            G[4] += -e3 * (B2 + e1 * e1 * B3) + e3 * R1;
            G[5] += e1 * (B2 + e3 * e3 * B3);
            G[6] += -e0 * (B1 + e3 * e3 * B3) - e0 * R1;
            G[7] += -e1 * (B1 + e3 * e3 * B3) - e1 * R1;
            G[8] += e3 * (B1 + 2.0 * B2 + e3 * e3 * B3);
        }
    } else {
        double Q = d0 * d0 + d1 * d1 + d2 * d2 + e2;
        double rs = rsqrt(Q);
        double z = alpha * (Q * rs);
        double f1, f2;
        if (z < 0.5) {
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int n = NTERMS - 1; n >= 0; --n) {
                s1 = s1 * z + kSeriesA.c1[n];
                s2 = s2 * z + kSeriesA.c2[n];
            }
            f1 = s1;
            f2 = s2;
        } else {
            double E = exp(-z);
            double iz = rs * inv_alpha;
            double iz2 = iz * iz;
            f1 = E * (1.0 + iz + iz2) - iz2;
            f2 = 3.0 * iz2 - E * (1.0 + 3.0 * iz + 3.0 * iz2);
        }
        double h2 = f2 * (rs * rs * rs);
        double H1 = f1 * rs + e2 * h2;
        G[0] = H1 + h2 * d0 * d0;
        G[1] = h2 * d0 * d1;
        G[2] = h2 * d0 * d2;
        G[3] = G[1];
        G[4] = H1 + h2 * d1 * d1;
        G[5] = h2 * d1 * d2;
        G[6] = G[2];
        G[7] = G[5];
        G[8] = H1 + h2 * d2 * d2;
    }
}

struct AsmArgs {
    const double* tgt;
    const double* src;
    const double* eps;
    int64_t M, N, eps_stride;
    double alpha, inv_alpha;
    double w_re, w_im;
    double* K64;
    int64_t ld;
    float* K32;
    int64_t ld32;
    double* Kc;  // interleaved complex, leading dimension ld (complex elements)
};

template <int KIND>
__global__ __launch_bounds__(WG) void assemble_kernel(AsmArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * RB;
    if (j >= a.N) return;
    const double sx = a.src[j], sy = a.src[a.N + j], sz = a.src[2 * a.N + j];
    const double e = a.eps[j * a.eps_stride];
    const double e2 = e * e;
#pragma unroll
    for (int t = 0; t < RB; ++t) {
        const int64_t i = i0 + t;
        if (i >= a.M) break;
        double G[9];
        block3x3<KIND>(a.tgt[i], a.tgt[a.M + i], a.tgt[2 * a.M + i], sx, sy, sz, e2, a.alpha, a.inv_alpha, G);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int64_t row = 3 * i + r;
            if (a.K64) {
                double* p = a.K64 + row * a.ld + 3 * j;
                p[0] = a.w_re * G[3 * r];
                p[1] = a.w_re * G[3 * r + 1];
                p[2] = a.w_re * G[3 * r + 2];
            }
            if (a.K32) {
                float* q = a.K32 + row * a.ld32 + 3 * j;
                q[0] = (float)(a.w_re * G[3 * r]);
                q[1] = (float)(a.w_re * G[3 * r + 1]);
                q[2] = (float)(a.w_re * G[3 * r + 2]);
            }
            if (a.Kc) {
                double2* c = reinterpret_cast<double2*>(a.Kc) + row * a.ld + 3 * j;
#pragma unroll
                for (int b = 0; b < 3; ++b) c[b] = make_double2(a.w_re * G[3 * r + b], a.w_im * G[3 * r + b]);
            }
        }
    }
}

struct NearArgs {
    const double* tgt;
    const double* quad;
    const double* eps;
    const int64_t* order;
    const int64_t* start;
    int64_t M, Q, n_force, eps_stride;
    double alpha, inv_alpha, w;
    double* K64;
    int64_t ld;
    float* K32;
    int64_t ld32;
};

// One lane per (target, force point): the in-order sum over that force
// point's quadrature group (core/assemble.py:51-58 accumulates the same
// entries in increasing k).
template <int KIND>
__global__ __launch_bounds__(WG) void assemble_nearest_kernel(NearArgs a)
{
    const int64_t j = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * RB;
    if (j >= a.n_force) return;
    const int64_t kb = a.start[j], ke = a.start[j + 1];
#pragma unroll 1
    for (int t = 0; t < RB; ++t) {
        const int64_t i = i0 + t;
        if (i >= a.M) break;
        const double tx = a.tgt[i], ty = a.tgt[a.M + i], tz = a.tgt[2 * a.M + i];
        double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t p = kb; p < ke; ++p) {
            const int64_t k = a.order[p];
            const double e = a.eps[k * a.eps_stride];
            double G[9];
            block3x3<KIND>(tx, ty, tz, a.quad[k], a.quad[a.Q + k], a.quad[2 * a.Q + k], e * e, a.alpha, a.inv_alpha,
                           G);
#pragma unroll
            for (int q = 0; q < 9; ++q) S[q] += G[q];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int64_t row = 3 * i + r;
            if (a.K64) {
                double* p = a.K64 + row * a.ld + 3 * j;
                p[0] = a.w * S[3 * r];
                p[1] = a.w * S[3 * r + 1];
                p[2] = a.w * S[3 * r + 2];
            }
            if (a.K32) {
                float* q = a.K32 + row * a.ld32 + 3 * j;
                q[0] = (float)(a.w * S[3 * r]);
                q[1] = (float)(a.w * S[3 * r + 1]);
                q[2] = (float)(a.w * S[3 * r + 2]);
            }
        }
    }
}

// Nearest force point of every quadrature point: one lane per quadrature
// point, force points staged through LDS in tiles of 256, squared distances in
// FP64, the lowest index kept on ties (replaces the cKDTree query of
// core/assemble.py:61-65).
__global__ __launch_bounds__(WG) void nearest_kernel(int64_t Q, int64_t N, const double* __restrict__ quad,
                                                     const double* __restrict__ force, int64_t* __restrict__ out)
{
    __shared__ double sx[WG], sy[WG], sz[WG];
    const int64_t k = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t kc = k < Q ? k : Q - 1;
    const double qx = quad[kc], qy = quad[Q + kc], qz = quad[2 * Q + kc];
    double best = 1.0 / 0.0;
    int64_t arg = 0;
    for (int64_t j0 = 0; j0 < N; j0 += WG) {
        const int64_t j = j0 + threadIdx.x;
        __syncthreads();
        if (j < N) {
            sx[threadIdx.x] = force[j];
            sy[threadIdx.x] = force[N + j];
            sz[threadIdx.x] = force[2 * N + j];
        }
        __syncthreads();
        const int nt = (int)min((int64_t)WG, N - j0);
        for (int s = 0; s < nt; ++s) {
            double dx = qx - sx[s], dy = qy - sy[s], dz = qz - sz[s];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) {
                best = d2;
                arg = j0 + s;
            }
        }
    }
    if (k < Q) out[k] = arg;
}

// Diagonal blocks of the resistance matrix for bodies of P consecutive points:
// KB[b][3i + a][3j + c] = w G_ac(x_{bP+i}, x_{bP+j}; eps_{bP+j}), i, j < P,
// each block 3P x 3P row-major (the block-Jacobi preconditioner of GMRES).
// One lane per (block, target, source); consecutive lanes write consecutive
// sources of a row.
template <int KIND>
__global__ __launch_bounds__(WG) void assemble_blocks_kernel(const double* X, int64_t N, int P, int64_t nblk,
                                                             const double* eps, int64_t eps_stride, double alpha,
                                                             double inv_alpha, double w, double* KB)
{
    const int64_t t = (int64_t)blockIdx.x * WG + threadIdx.x;
    const int64_t PP = (int64_t)P * P;
    const int64_t b = t / PP;
    if (b >= nblk) return;
    const int rem = (int)(t % PP), i = rem / P, j = rem % P;
    const int64_t gi = b * P + i, gj = b * P + j;
    const double e = eps[gj * eps_stride];
    double G[9];
    block3x3<KIND>(X[gi], X[N + gi], X[2 * N + gi], X[gj], X[N + gj], X[2 * N + gj], e * e, alpha, inv_alpha, G);
    const int64_t m = 3 * (int64_t)P;
    double* K = KB + b * m * m;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double* row = K + (3 * i + a) * m + 3 * j;
        row[0] = w * G[3 * a];
        row[1] = w * G[3 * a + 1];
        row[2] = w * G[3 * a + 2];
    }
}

}  // namespace

int nearest_launch(int64_t Q, int64_t N, const double* quad, const double* force, int64_t* out, hipStream_t stream)
{
    FL_CHECK(Q >= 0 && N >= 0, "negative size");
    FL_CHECK(N > 0 || Q == 0, "no force points");
    if (Q == 0) return FL_OK;
    FL_CHECK(quad && force && out, "null pointer argument");
    hipLaunchKernelGGL(nearest_kernel, dim3((unsigned)((Q + WG - 1) / WG)), dim3(WG), 0, stream, Q, N, quad, force,
                       out);
    FL_LAUNCHED();
    return FL_OK;
}

int assemble_launch(int kind, double alpha, int64_t M, int64_t N, const double* tgt, const double* src,
                    const double* eps, int64_t eps_stride, double mu_re, double mu_im, double* K64, int64_t ld,
                    float* K32, int64_t ld32, double* Kc, hipStream_t stream)
{
    FL_CHECK(kind >= 0 && kind <= 2, "unknown kernel kind %d", kind);
    FL_CHECK(M >= 0 && N >= 0, "negative size");
    FL_CHECK(eps_stride == 0 || eps_stride == 1, "eps_stride must be 0 or 1");
    FL_CHECK(K64 || K32 || Kc, "no output requested");
    FL_CHECK(!(K64 || K32) || mu_im == 0.0, "real outputs need a real mu");
    FL_CHECK(!K64 || ld >= 3 * N, "ld too small");
    FL_CHECK(!K32 || ld32 >= 3 * N, "ld32 too small");
    FL_CHECK(!Kc || ld >= 3 * N, "ld too small");
    FL_CHECK(mu_re != 0.0 || mu_im != 0.0, "mu must be nonzero");
    if (M == 0 || N == 0) return FL_OK;
    AsmArgs a;
    a.tgt = tgt;
    a.src = src;
    a.eps = eps;
    a.M = M;
    a.N = N;
    a.eps_stride = eps_stride;
    a.alpha = alpha;
    a.inv_alpha = alpha > 0 ? 1.0 / alpha : 0.0;
    const double s = kernel_scale(kind);
    const double den = mu_re * mu_re + mu_im * mu_im;
    a.w_re = s * mu_re / den;
    a.w_im = -s * mu_im / den;
    if (mu_im == 0.0) a.w_re = s / mu_re;
    a.K64 = K64;
    a.ld = ld;
    a.K32 = K32;
    a.ld32 = ld32;
    a.Kc = Kc;
    dim3 grid((unsigned)((N + WG - 1) / WG), (unsigned)((M + RB - 1) / RB));
    if (kind == FL_STOKESLET3D) hipLaunchKernelGGL(assemble_kernel<FL_STOKESLET3D>, grid, dim3(WG), 0, stream, a);
    else if (kind == FL_WALL_STOKESLET3D)
        hipLaunchKernelGGL(assemble_kernel<FL_WALL_STOKESLET3D>, grid, dim3(WG), 0, stream, a);
    else hipLaunchKernelGGL(assemble_kernel<FL_BRINKMANLET3D>, grid, dim3(WG), 0, stream, a);
    FL_LAUNCHED();
    return FL_OK;
}

int assemble_nearest_launch(int kind, double alpha, int64_t M, int64_t Q, int64_t n_force, const double* tgt,
                            const double* quad, const double* eps, int64_t eps_stride, const int64_t* order,
                            const int64_t* start, double mu_re, double* K64, int64_t ld, float* K32, int64_t ld32,
                            hipStream_t stream)
{
    FL_CHECK(kind >= 0 && kind <= 2, "unknown kernel kind %d", kind);
    FL_CHECK(M >= 0 && Q >= 0 && n_force >= 0, "negative size");
    FL_CHECK(eps_stride == 0 || eps_stride == 1, "eps_stride must be 0 or 1");
    FL_CHECK(K64 || K32, "K64 or K32 is required");
    FL_CHECK(!K64 || ld >= 3 * n_force, "ld too small");
    FL_CHECK(!K32 || ld32 >= 3 * n_force, "ld32 too small");
    FL_CHECK(mu_re != 0.0, "mu must be nonzero");
    if (M == 0 || n_force == 0) return FL_OK;
    NearArgs a;
    a.tgt = tgt;
    a.quad = quad;
    a.eps = eps;
    a.order = order;
    a.start = start;
    a.M = M;
    a.Q = Q;
    a.n_force = n_force;
    a.eps_stride = eps_stride;
    a.alpha = alpha;
    a.inv_alpha = alpha > 0 ? 1.0 / alpha : 0.0;
    a.w = kernel_scale(kind) / mu_re;
    a.K64 = K64;
    a.ld = ld;
    a.K32 = K32;
    a.ld32 = ld32;
    dim3 grid((unsigned)((n_force + WG - 1) / WG), (unsigned)((M + RB - 1) / RB));
    if (kind == FL_STOKESLET3D)
        hipLaunchKernelGGL(assemble_nearest_kernel<FL_STOKESLET3D>, grid, dim3(WG), 0, stream, a);
    else if (kind == FL_WALL_STOKESLET3D)
        hipLaunchKernelGGL(assemble_nearest_kernel<FL_WALL_STOKESLET3D>, grid, dim3(WG), 0, stream, a);
    else hipLaunchKernelGGL(assemble_nearest_kernel<FL_BRINKMANLET3D>, grid, dim3(WG), 0, stream, a);
    FL_LAUNCHED();
    return FL_OK;
}

int assemble_blocks_launch(int kind, double alpha, int64_t N, int P, const double* X, const double* eps,
                           int64_t eps_stride, double w, double* KB, hipStream_t stream)
{
    FL_CHECK(kind >= 0 && kind <= 2, "unknown kernel kind %d", kind);
    FL_CHECK(P >= 1 && N >= 0 && N % P == 0, "N must be a multiple of the block size");
    FL_CHECK(eps_stride == 0 || eps_stride == 1, "eps_stride must be 0 or 1");
    const int64_t nblk = N / P;
    const int64_t total = nblk * (int64_t)P * P;
    if (total == 0) return FL_OK;
    const dim3 grid((unsigned)((total + WG - 1) / WG));
    const double inv_alpha = alpha > 0 ? 1.0 / alpha : 0.0;
    if (kind == FL_STOKESLET3D)
        hipLaunchKernelGGL(assemble_blocks_kernel<FL_STOKESLET3D>, grid, dim3(WG), 0, stream, X, N, P, nblk, eps,
                           eps_stride, alpha, inv_alpha, w, KB);
    else if (kind == FL_WALL_STOKESLET3D)
        hipLaunchKernelGGL(assemble_blocks_kernel<FL_WALL_STOKESLET3D>, grid, dim3(WG), 0, stream, X, N, P, nblk, eps,
                           eps_stride, alpha, inv_alpha, w, KB);
    else hipLaunchKernelGGL(assemble_blocks_kernel<FL_BRINKMANLET3D>, grid, dim3(WG), 0, stream, X, N, P, nblk, eps,
                            eps_stride, alpha, inv_alpha, w, KB);
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace fl
