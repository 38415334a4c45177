// Blocked right-looking LU with partial pivoting, triangular solves and FP64
// iterative refinement, for the dense resistance matrices.
//
// Panel step k (width nb <= NB = 128), in row-major storage:
//   1. one workgroup factors the nb x nb diagonal block by GEPP in registers,
//      its pivot search restricted to the block's rows, and forms inv(L11)
//      and inv(U11), kept in the factor's aux buffer for the triangular solves;
//   2. L21 = A21 inv(U11) on MFMA, in place; the GEMM epilogue records
//      max |L21|.  If it is <= 1 every pivot chosen in step 1 is also the
//      pivot partial pivoting over the whole column would choose (all
//      multipliers below the block are bounded by the block's pivot), so the
//      factorization is exactly GEPP.  Otherwise the panel is restored from a
//      copy and refactored column by column with the pivot search over all
//      rows below the diagonal (the general path);
//   3. the panel's row interchanges are applied left and right of it;
//   4. U12 = inv(L11) A12 on MFMA, in place;
//   5. A22 -= L21 U12 on MFMA (FP32 or FP64 accumulation).
// The LAPACK pivot vector (ipiv[r] = row swapped with r) is also folded into a
// row permutation stored with the factor, so solves gather the right-hand
// side in one pass.
//
// Replaces LAPACK getrf/getrs behind scipy.linalg.solve
// (src/fluidlets/core/solve.py:66) and numpy.linalg.solve
// (src/fluidlets/swimming.py:101).
#include "fl_internal.h"
#include "mfma.h"


#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

namespace fl {

template <typename T>
int gemm_launch(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                T* C, int64_t ldc, unsigned int* absmax_bits, hipStream_t stream);
template <typename T>
int gemm_pair_launch(int64_t M1, int64_t N1, int64_t K1, T alpha1, const T* A1, int64_t lda1, const T* B1,
                     int64_t ldb1, T beta1, T* C1, int64_t ldc1, unsigned int* absmax1, int64_t M2, int64_t N2,
                     int64_t K2, T alpha2, const T* A2, int64_t lda2, const T* B2, int64_t ldb2, T beta2, T* C2,
                     int64_t ldc2, hipStream_t stream);

template <typename T>
int gemm_lower_update(int64_t n, int64_t K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc,
                      hipStream_t stream);

// split-bf16 trailing update (gemm_bf16x3.hip)
size_t split3_workspace_bytes(int64_t M, int64_t N, int64_t K);
int split3_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                  int64_t ldc, int lower, void* ws, size_t ws_bytes, hipStream_t st);
size_t f32_workspace_bytes(int64_t M, int64_t N, int64_t K);
int f32_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
               int64_t ldc, int lower, void* ws, size_t ws_bytes, hipStream_t st);

namespace {

constexpr int NB = 128;   // diagonal block (LDS-resident GEPP, inverted for the solves)
constexpr int NBO = 256;  // outer panel width: the trailing update runs with K = NBO
// Outer panel width with split-bf16 trailing products: the cheaper products
// leave the update bound by its passes over A22, which a wider panel halves.
// (Measured at n = 24,576: 256 -> 512 takes the mixed solve from 85 to 75 ms;
// 768 and 1024 gain 1 ms more but the factors' error grows with the width.)
constexpr int NBO_SPLIT = 512;

// ---------------------------------------------------------------- triangular inverses (general path)
// Used after a panel has been refactored column by column with full-column
// pivoting.  grid (2, NB / 16): blockIdx.x 0 forms X = inv(L11) (unit lower), 1 forms
// X = inv(U11) (upper); blockIdx.y selects 16 columns.  The factored block is
// staged in LDS; column j of X is computed by 8 consecutive lanes of one wave
// (each summing an eighth of the dot product, combined by shuffles), so rows
// are produced without workgroup barriers.  Output: nb x nb row-major (ld NB),
// zero outside the triangle.
constexpr int TR_COLS = 16;

template <typename T>
__global__ __launch_bounds__(128) void trtri_kernel(const T* A, int64_t lda, int nb, T* invL, T* invU)
{
    constexpr int LX = NB + 1;
    __shared__ T F[NB * LX];
    __shared__ T X[NB * TR_COLS];
    const int tid = threadIdx.x;
    const int jj = tid / 8, part = tid % 8;
    const int j = blockIdx.y * TR_COLS + jj;
    for (int idx = tid; idx < nb * nb; idx += 128) F[(idx / nb) * LX + idx % nb] = A[(idx / nb) * lda + idx % nb];
    __syncthreads();
    T* out = blockIdx.x == 0 ? invL : invU;
    if (blockIdx.x == 0) {
        for (int i = 0; i < nb; ++i) {
            T s = T(0);
            // sum over kk in [j, i) of L[i][kk] X[kk][j]
            for (int kk = j + part; kk < i; kk += 8) s += F[i * LX + kk] * X[kk * TR_COLS + jj];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            const T v = i > j ? -s : (i == j ? T(1) : T(0));
            if (part == 0) X[i * TR_COLS + jj] = v;
        }
    } else {
        for (int i = nb - 1; i >= 0; --i) {
            T s = T(0);
            // sum over kk in (i, j] of U[i][kk] X[kk][j]
            for (int kk = i + 1 + part; kk <= j && kk < nb; kk += 8) s += F[i * LX + kk] * X[kk * TR_COLS + jj];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            s += __shfl_xor(s, 4);
            const T d = F[i * LX + i];
            const T v = i < j ? -s / d : (i == j ? T(1) / d : T(0));
            if (part == 0) X[i * TR_COLS + jj] = v;
        }
    }
    __syncthreads();
    for (int idx = tid; idx < nb * TR_COLS; idx += 128) {
        const int i = idx / TR_COLS, c = blockIdx.y * TR_COLS + idx % TR_COLS;
        if (c < nb) out[i * NB + c] = X[idx];
    }
}

// ---------------------------------------------------------------- register-resident diagonal block
// One workgroup of 256 lanes factors the nb x nb diagonal block (nb <= NB) by
// GEPP and forms inv(L11) and inv(U11) for the panel's products and the
// triangular solves.
//
// Factorization: lane (slot s = tid >> 1, half h = tid & 1) holds the slot's
// entries in columns 2k + h, k < 64, in registers.  Rows are not moved: each
// slot carries its position pos in LAPACK's row order after the interchanges
// so far.  Step c takes the active slot with the largest |a(c)| (the smallest
// pos on ties, as isamax does), records ipiv[c] = its pos, exchanges its
// position with the slot at position c, and eliminates column c from the
// other active slots (multiplier a(c) / pivot, as getf2 scales by the
// reciprocal) with the pivot row read from LDS.  Each wave publishes its best
// candidate and that candidate's row, double-buffered by column parity, so a
// column costs one barrier.  Slot s is finally written to row pos(s), which
// is the row-interchanged LU factor.
//
// Inverses: each of the four 64 x 64 diagonal triangles is inverted by
// substitution (one wave per triangle, one lane per column); the off-diagonal
// blocks follow from inv(L)21 = -inv(L22) L21 inv(L11) and
// inv(U)12 = -inv(U11) U12 inv(U22).  Outputs are NB x NB, ld NB, zero outside
// the triangle (identity beyond nb).
// Lane exchange through DPP (a VALU modifier, no LDS round trip).  CTRL:
// quad_perm [1,0,3,2] = 0xB1 (swap neighbours), quad_perm [2,3,0,1] = 0x4E,
// row_ror:4 = 0x124, row_ror:8 = 0x128 (rotations inside 16-lane rows).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// 1 / x: the hardware reciprocal in FP32 (1 ulp; the FP32 factors only
// precondition FP64 refinement), the correctly rounded division in FP64.
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

__device__ __forceinline__ float read_lane(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

__device__ __forceinline__ double read_lane(double x, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                            __builtin_amdgcn_readlane(__double2loint(x), l));
}

// Largest v over the wave (every lane receives it).
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    v = fmax(v, dpp_mov<0xB1>(v));
    v = fmax(v, dpp_mov<0x4E>(v));
    v = fmax(v, dpp_mov<0x124>(v));
    v = fmax(v, dpp_mov<0x128>(v));
    return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}

// 16-byte vector of T for LDS and global moves.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};

// Column j of the inverse of the 64 x 64 unit lower triangle of F (pitch lf),
// by forward substitution, into x[i] (row i).
template <typename T>
__device__ __forceinline__ void tri_inverse_lower(const T* F, int lf, int j, T (&x)[64])
{
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        T s0 = T(0), s1 = T(0);
#pragma unroll
        for (int m = 0; m + 1 < 64; m += 2) {
            if (m + 1 < i) {
                s0 = fma(F[i * lf + m], x[m], s0);
                s1 = fma(F[i * lf + m + 1], x[m + 1], s1);
            }
        }
        if (i & 1) s0 = fma(F[i * lf + i - 1], x[i - 1], s0);
        x[i] = i < j ? T(0) : (i == j ? T(1) : -(s0 + s1));
    }
}

// Column j of the inverse of the 64 x 64 upper triangle of F, by backward
// substitution; x[q] holds row 63 - q.
template <typename T>
__device__ __forceinline__ void tri_inverse_upper(const T* F, int lf, int j, T (&x)[64])
{
#pragma unroll
    for (int q = 0; q < 64; ++q) {
        const int i = 63 - q;
        T s0 = T(0), s1 = T(0);
#pragma unroll
        for (int m = 0; m + 1 < 64; m += 2) {
            if (m + 1 < q) {
                s0 = fma(F[i * lf + 63 - m], x[m], s0);
                s1 = fma(F[i * lf + 62 - m], x[m + 1], s1);
            }
        }
        if (q & 1) s0 = fma(F[i * lf + 64 - q], x[q - 1], s0);
        const T rd = T(1) / F[i * lf + i];
        x[q] = i > j ? T(0) : (i == j ? rd : -(s0 + s1) * rd);
    }
}

// The off-diagonal blocks of inv(L) and inv(U) on MFMA, from the packed
// diagonal inverses in S (block dd: inv(L_dd) strictly below its diagonal,
// 1 implied on it; inv(U_dd) on and above it):
//   phase 1: T1 = L21 inv(L11), T2 = U12 inv(U22), written over L21 and U12;
//   phase 2: inv(L)21 = -inv(L22) T1, inv(U)12 = -inv(U11) T2, to invL, invU.
// Waves 0, 1 form the L products, waves 2, 3 the U products, each half of
// the 64 x 64 output in MFMA blocks (32 x 32 FP32, 16 x 16 FP64, K = 64).
template <typename T>
__device__ __forceinline__ void block_inverse_products(T* S, T* invL, T* invU, int wave, int lane)
{
    using M = mfma::Mma<T>;
    constexpr int H = NB / 2, LS = NB + 1, TB = M::B, TPE = H / TB, TW = TPE * TPE / 2;
    const bool lside = wave < 2;
    const int t0 = (wave & 1) * TW;
    typename M::acc_t acc[TW];
#pragma unroll
    for (int u = 0; u < TW; ++u) {
        const int t = t0 + u, ar = (t / TPE) * TB + M::a_row(lane), bc = (t % TPE) * TB + M::d_col(lane);
        acc[u] = M::zero();
#pragma unroll 4
        for (int k0 = 0; k0 < H; k0 += M::KS) {
            const int k = k0 + M::a_k(lane);
            T a, b;
            if (lside) {
                a = S[(H + ar) * LS + k];                                      // L21[ar][k]
                b = k > bc ? S[k * LS + bc] : (k == bc ? T(1) : T(0));         // inv(L11)[k][bc]
            } else {
                a = S[ar * LS + H + k];                                        // U12[ar][k]
                b = k <= bc ? S[(H + k) * LS + H + bc] : T(0);                 // inv(U22)[k][bc]
            }
            acc[u] = M::mma(a, b, acc[u]);
        }
    }
    __syncthreads();  // every read of L21 and U12 is done
#pragma unroll
    for (int u = 0; u < TW; ++u) {
        const int t = t0 + u, r0 = (t / TPE) * TB, col = (t % TPE) * TB + M::d_col(lane);
#pragma unroll
        for (int i = 0; i < M::NACC; ++i) {
            const int row = r0 + M::d_row(lane, i);
            if (lside) S[(H + row) * LS + col] = acc[u][i];
            else S[row * LS + H + col] = acc[u][i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TW; ++u) {
        const int t = t0 + u, ar = (t / TPE) * TB + M::a_row(lane), bc = (t % TPE) * TB + M::d_col(lane);
        acc[u] = M::zero();
#pragma unroll 4
        for (int k0 = 0; k0 < H; k0 += M::KS) {
            const int k = k0 + M::a_k(lane);
            T a, b;
            if (lside) {
                a = k < ar ? S[(H + ar) * LS + H + k] : (k == ar ? T(1) : T(0));  // inv(L22)[ar][k]
                b = S[(H + k) * LS + bc];                                      // T1[k][bc]
            } else {
                a = k >= ar ? S[ar * LS + k] : T(0);                           // inv(U11)[ar][k]
                b = S[k * LS + H + bc];                                        // T2[k][bc]
            }
            acc[u] = M::mma(a, b, acc[u]);
        }
        const int r0 = (t / TPE) * TB, col = (t % TPE) * TB + M::d_col(lane);
#pragma unroll
        for (int i = 0; i < M::NACC; ++i) {
            const int row = r0 + M::d_row(lane, i);
            if (lside) invL[(H + row) * NB + col] = -acc[u][i];
            else invU[row * NB + H + col] = -acc[u][i];
        }
    }
}

// inv(L) and inv(U) of the factored NB x NB block in S (pitch NB + 1), as
// NB x NB row-major arrays (ld NB, zero outside the triangle).  The four
// 64 x 64 diagonal triangles are inverted by substitution (waves 0, 1: L11,
// L22; waves 2, 3: U11, U22; one lane per column, in registers) and packed
// back into S's diagonal blocks once every wave has read its triangle: block
// dd holds inv(L_dd) strictly below its diagonal and inv(U_dd) on and above
// it.  The off-diagonal blocks follow on MFMA (block_inverse_products),
// overwriting S's L21 and U12.  All 256 lanes call it.
template <typename T>
__device__ __forceinline__ void block_inverses(T* S, T* invL, T* invU, int tid)
{
    constexpr int H = NB / 2;
    constexpr int LS = NB + 1;
    const int lane = tid & 63, wave = tid >> 6;
    const int d = (wave & 1) * H;  // this wave's diagonal block
    T x[64];
    if (wave < 2) tri_inverse_lower<T>(S + d * (LS + 1), LS, lane, x);
    else tri_inverse_upper<T>(S + d * (LS + 1), LS, lane, x);
    __syncthreads();  // every triangle of S has been read
    if (wave < 2) {
#pragma unroll
        for (int i = 0; i < 64; ++i) {
            if (i > lane) S[(d + i) * LS + d + lane] = x[i];
            invL[(d + i) * NB + d + lane] = x[i];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            const int i = 63 - q;
            if (i <= lane) S[(d + i) * LS + d + lane] = x[q];
            invU[(d + i) * NB + d + lane] = x[q];
        }
    }
    if (wave == 0)
#pragma unroll
        for (int i = 0; i < 64; ++i) invL[i * NB + 64 + lane] = T(0);
    if (wave == 3)
#pragma unroll
        for (int i = 0; i < 64; ++i) invU[(64 + i) * NB + lane] = T(0);
    __syncthreads();  // the packed inverses are read by every wave
    block_inverse_products<T>(S, invL, invU, wave, lane);
}

template <typename T>
__global__ __launch_bounds__(256) void diag_lu_kernel(T* A, int64_t lda, int nb, int64_t k0, int32_t* ipiv,
                                                      int32_t* info, T* invL, T* invU)
{
    constexpr int H = NB / 2;
    constexpr int LS = NB + 1;
    constexpr int GC = 8;  // columns per unrolled group; the register window advances by GC / 2
    constexpr int V = Vec16<T>::n;
    __shared__ T S[NB * LS];
    __shared__ __attribute__((aligned(16))) T prow[2][4][NB];  // per-wave candidate rows, by column parity
    __shared__ T cand_v[2][4];
    __shared__ int cand_p[2][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slot = tid >> 1, h = tid & 1;

    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx / NB, c = idx % NB;
        S[r * LS + c] = (r < nb && c < nb) ? A[r * lda + c] : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    // a[k] holds column 2 (k + kb) + h of this slot, kb = GC / 2 * group (a window
    // that drops the finished columns into S row `slot` as the factorization advances)
    T a[H];
#pragma unroll
    for (int k = 0; k < H; ++k) a[k] = S[slot * LS + 2 * k + h];
    __syncthreads();

    int pos = slot;
    bool done = slot >= nb;
    for (int g = 0; g < NB / GC; ++g) {
#pragma unroll
        for (int cl = 0; cl < GC; ++cl) {
            const int c = g * GC + cl;
            if (c < nb) {  // uniform
                const int kc = cl >> 1, hc = cl & 1, buf = cl & 1;
                const bool cand = h == hc && !done;
                const T v = cand ? (a[kc] < T(0) ? -a[kc] : a[kc]) : T(-1);
                // wave candidate: largest |a(c)|, smallest position among equals
                const T wv = wave_max(v);
                unsigned long long tied = __ballot(cand && v == wv);
                int wp = 0x7fffffff;
                while (tied) {  // uniform; one pass unless values tie exactly
                    const int l = __ffsll((long long)tied) - 1;
                    wp = min(wp, __builtin_amdgcn_readlane(pos, l));
                    tied &= tied - 1;
                }
                if (!done && pos == wp) {
                    auto* dst = reinterpret_cast<typename Vec16<T>::type*>(&prow[buf][wave][h * H]);
#pragma unroll
                    for (int k = 0; k < H; k += V) {
                        typename Vec16<T>::type u;
                        T* up = reinterpret_cast<T*>(&u);
#pragma unroll
                        for (int q = 0; q < V; ++q) up[q] = a[k + q];
                        dst[k / V] = u;
                    }
                }
                if (lane == 0) {
                    cand_v[buf][wave] = wv;
                    cand_p[buf][wave] = wp;
                }
                __syncthreads();
                T bv = cand_v[buf][0];
                int bp = cand_p[buf][0], bw = 0;
#pragma unroll
                for (int w = 1; w < 4; ++w) {
                    const T ov = cand_v[buf][w];
                    const int op = cand_p[buf][w];
                    if (ov > bv || (ov == bv && op < bp)) {
                        bv = ov;
                        bp = op;
                        bw = w;
                    }
                }
                const T pv = prow[buf][bw][hc * H + kc];
                if (tid == 0) {
                    ipiv[k0 + c] = (int32_t)(k0 + bp);
                    if (pv == T(0)) atomicCAS(info, 0, (int32_t)(k0 + c + 1));
                }
                const T rp = pv != T(0) ? T(1) / pv : T(1);
                const bool piv = !done && pos == bp;
                if (pos == c) pos = bp;
                if (piv) pos = c;
                done = done || piv;
                const T lm = a[kc] * rp;
                const T lo = dpp_mov<0xB1>(lm);
                const T l = done ? T(0) : (h == hc ? lm : lo);
                const auto* ph = reinterpret_cast<const typename Vec16<T>::type*>(&prow[buf][bw][h * H]);
                const T akc = a[kc];
#pragma unroll
                for (int k0v = 0; k0v < H; k0v += V) {
                    if (k0v + V - 1 >= kc) {
                        const typename Vec16<T>::type u = ph[k0v / V];
                        const T* up = reinterpret_cast<const T*>(&u);
#pragma unroll
                        for (int q = 0; q < V; ++q)
                            if (k0v + q > kc) a[k0v + q] = fma(-l, up[q], a[k0v + q]);
                        if (kc >= k0v && kc < k0v + V) {
                            // column 2 kc + h: the multiplier (h == hc) or an update if right of c
                            const T upd = fma(-l, up[kc - k0v], akc);
                            a[kc] = h == hc ? (done ? akc : lm) : (2 * kc + h > cl ? upd : akc);
                        }
                    }
                }
            }
        }
        // columns 2 (kb + k) + h, k < GC / 2, are final: store them and advance the window
#pragma unroll
        for (int k = 0; k < GC / 2; ++k) S[slot * LS + 2 * (g * (GC / 2) + k) + h] = a[k];
#pragma unroll
        for (int k = 0; k < H - GC / 2; ++k) a[k] = a[k + GC / 2];
#pragma unroll
        for (int k = H - GC / 2; k < H; ++k) a[k] = T(0);
    }
    // rows by slot -> rows by position
    __syncthreads();
#pragma unroll
    for (int k = 0; k < H; ++k) a[k] = S[slot * LS + 2 * k + h];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < H; ++k) S[pos * LS + 2 * k + h] = a[k];
    __syncthreads();
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        A[r * lda + c] = S[r * LS + c];
    }

    block_inverses<T>(S, invL, invU, tid);
}

// S (pitch NB + 1) = the nb x nb block at A, padded to NB x NB with the
// identity.  A full block with 16-byte aligned rows is read as 16-byte
// vectors, in two rounds whose loads are all issued before any LDS store.
template <typename T>
__device__ __forceinline__ void load_diag_block(const T* A, int64_t lda, int nb, T* S, int tid)
{
    constexpr int LS = NB + 1;
    constexpr int V = Vec16<T>::n;
    constexpr int RV = NB / V;                 // vectors per row
    constexpr int PER = NB * RV / 256 / 2;     // vectors per lane per round
    using VT = typename Vec16<T>::type;
    if (nb == NB && lda % V == 0 && ((uintptr_t)A % 16) == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            VT buf[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = (h * PER + k) * 256 + tid;
                buf[k] = *reinterpret_cast<const VT*>(A + (int64_t)(idx / RV) * lda + (idx % RV) * V);
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = (h * PER + k) * 256 + tid;
                const T* e = reinterpret_cast<const T*>(&buf[k]);
#pragma unroll
                for (int q = 0; q < V; ++q) S[(idx / RV) * LS + (idx % RV) * V + q] = e[q];
            }
        }
    } else {
        for (int idx = tid; idx < NB * NB; idx += 256) {
            const int r = idx / NB, c = idx % NB;
            S[r * LS + c] = (r < nb && c < nb) ? A[r * lda + c] : (r == c ? T(1) : T(0));
        }
    }
}

// ---------------------------------------------------------------- diagonal block without interchanges
// The resistance matrices never need a row interchange (every pivot is the
// largest entry left in its column), and then GEPP is LU without
// interchanges.  This kernel factors the nb x nb diagonal block that way, by
// 32 x 32 sub-blocks (right-looking):
//   leaf   wave 0 factors the 32 x 32 diagonal sub-block, one row per lane,
//          the pivot row broadcast with readlane (no barrier per column);
//   panels rows below: L = A inv(U_leaf), columns right: U = inv(L_leaf) A,
//          by substitution, one row or column per lane;
//   update the trailing sub-blocks, 4 x 4 register tiles per lane;
// records ipiv[c] = c and folds the largest |multiplier| into *lmax_bits
// (float bits, atomic max; infinite for a zero pivot).  If that exceeds 1,
// some pivot was not the largest in its column and the caller refactors the
// panel with interchanges.
// Then forms inv(L11), inv(U11) as the pivoting kernel does.
constexpr int SUBB = 32;

template <typename T>
__global__ __launch_bounds__(256) void diag_nopiv_kernel(T* A, int64_t lda, int nb, int64_t k0, int32_t* ipiv,
                                                         unsigned int* lmax_bits, T* invL, T* invU)
{
    constexpr int LS = NB + 1;
    __shared__ T S[NB * LS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    load_diag_block<T>(A, lda, nb, S, tid);
    for (int c = tid; c < nb; c += 256) ipiv[k0 + c] = (int32_t)(k0 + c);
    __syncthreads();
    T lmax = T(0);
    for (int b = 0; b < NB / SUBB; ++b) {
        const int o = b * SUBB;
        if (o >= nb) break;  // uniform: the rest is identity padding
        // leaf: rows o .. o + 31 by lanes 0 .. 31 of wave 0
        if (wave == 0) {
            T a[SUBB];
            const int r = o + (lane & (SUBB - 1));
#pragma unroll
            for (int j = 0; j < SUBB; ++j) a[j] = S[r * LS + o + j];
#pragma unroll
            for (int c = 0; c < SUBB; ++c) {
                const T pv = read_lane(a[c], c);
                if (pv == T(0)) lmax = T(INFINITY);  // a zero pivot: leave it to the pivoting tiers
                const T rp = pv != T(0) ? T(1) / pv : T(1);
                const bool below = (lane & (SUBB - 1)) > c && lane < SUBB;
                const T l = below ? a[c] * rp : T(0);
                if (below) {
                    a[c] = l;
                    lmax = fmax(lmax, l < T(0) ? -l : l);
                }
#pragma unroll
                for (int j = c + 1; j < SUBB; ++j) a[j] = fma(-l, read_lane(a[j], c), a[j]);
            }
            if (lane < SUBB)
#pragma unroll
                for (int j = 0; j < SUBB; ++j) S[r * LS + o + j] = a[j];
        }
        __syncthreads();
        const int rest = NB - o - SUBB;
        if (rest > 0) {
            if (tid < 128) {
                // row i below the leaf: x U_leaf = a
                const int i = o + SUBB + tid;
                if (tid < rest) {
                    T x[SUBB];
#pragma unroll
                    for (int j = 0; j < SUBB; ++j) x[j] = S[i * LS + o + j];
#pragma unroll
                    for (int c = 0; c < SUBB; ++c) {
                        const T d = S[(o + c) * LS + o + c];
                        if (d == T(0)) lmax = T(INFINITY);
                        x[c] = x[c] * (d != T(0) ? T(1) / d : T(1));
                        lmax = fmax(lmax, x[c] < T(0) ? -x[c] : x[c]);
#pragma unroll
                        for (int j = c + 1; j < SUBB; ++j) x[j] = fma(-x[c], S[(o + c) * LS + o + j], x[j]);
                    }
#pragma unroll
                    for (int j = 0; j < SUBB; ++j) S[i * LS + o + j] = x[j];
                }
            } else {
                // column j right of the leaf: L_leaf y = a (unit diagonal)
                const int jc = o + SUBB + (tid - 128);
                if (tid - 128 < rest) {
                    T y[SUBB];
#pragma unroll
                    for (int i = 0; i < SUBB; ++i) y[i] = S[(o + i) * LS + jc];
#pragma unroll
                    for (int c = 0; c < SUBB; ++c)
#pragma unroll
                        for (int i = c + 1; i < SUBB; ++i) y[i] = fma(-S[(o + i) * LS + o + c], y[c], y[i]);
#pragma unroll
                    for (int i = 0; i < SUBB; ++i) S[(o + i) * LS + jc] = y[i];
                }
            }
            __syncthreads();
            // trailing update of rows and columns o + 32 .. NB - 1, 4 x 4 per tile
            const int q4 = rest / 4;
            for (int tile = tid; tile < q4 * q4; tile += 256) {
                const int i0 = o + SUBB + (tile / q4) * 4, j0 = o + SUBB + (tile % q4) * 4;
                T acc[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] = S[(i0 + u) * LS + j0 + v];
#pragma unroll 4
                for (int k = 0; k < SUBB; ++k) {
                    T lv[4], uv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) lv[u] = S[(i0 + u) * LS + o + k];
#pragma unroll
                    for (int v = 0; v < 4; ++v) uv[v] = S[(o + k) * LS + j0 + v];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] = fma(-lv[u], uv[v], acc[u][v]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) S[(i0 + u) * LS + j0 + v] = acc[u][v];
            }
            __syncthreads();
        }
    }
    // largest multiplier of the block -> *lmax_bits
    const T wl = wave_max(lmax);
    if (lane == 0) atomicMax(lmax_bits, __float_as_uint((float)wl));
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx / nb, c = idx % nb;
        A[r * lda + c] = S[r * LS + c];
    }
    block_inverses<T>(S, invL, invU, tid);
}

// ---------------------------------------------------------------- diagonal block, products on MFMA
// The same factorization without interchanges as diag_nopiv_kernel, and
// the same outputs (factors, identity ipiv, largest |multiplier| folded into
// *lmax_bits, inv(L11) and inv(U11)), with every product on MFMA across the
// four waves.  The 128 x 128 block is a 4 x 4 grid of 32 x 32 sub-blocks;
// for kb = 0 .. 3:
//   leaf    wave 0 factors sub-block (kb, kb), one row per lane, the pivot
//           row broadcast with readlane; then the four waves invert the
//           16 x 16 diagonal triangles of its factors (one column per lane,
//           by substitution) and waves 0, 1 form the off-diagonal 16 x 16
//           blocks on MFMA: IL_kb = inv(L_kk), IU_kb = inv(U_kk);
//   panels  L_ib = A_ib IU_kb (i > kb) and U_kj = IL_kb A_kj (j > kb), one
//           32 x 32 x 32 MFMA product per wave;
//   update  A_ij -= L_ik U_kj for i, j > kb.
// The off-diagonal blocks of inv(L) and inv(U) follow block row by block
// row (column by column for inv(U)), from L X = I and Y U = I:
//   X_ij = -IL_i sum_{m = j}^{i - 1} L_im X_mj,  Y_ji = -(sum_{m = j}^{i - 1} Y_jm U_mi) IU_i,
// with X_mm = IL_m, Y_mm = IU_m, so row i needs only rows above it.
// Schedule: wave 0 carries the chain (update of diagonal block kb + 1 by
// step kb, leaf LU kb + 1); while it factors a leaf, waves 1 .. 3 do the
// other trailing updates of step kb and block row kb of X, Y; at the last
// leaf they also form the sums of row 3 that need no IL_3 and store the
// finished factor blocks.
//
// One wave's 32 x 32 MFMA accumulator (FP32: one v_mfma_f32_32x32x2_f32
// block; FP64: 2 x 2 v_mfma_f64_16x16x4_f64 blocks).
template <typename T>
struct Acc32;

template <>
struct Acc32<float> {
    mfma::floatx16 c;
    static __device__ int row(int lane, int i) { return 8 * (i / 4) + 4 * (lane / 32) + (i % 4); }
    static __device__ int col(int lane) { return lane & 31; }
    __device__ void zero() { c = mfma::floatx16{0}; }
    // c += s A B over K = 32: A (32 x 32, pitch la), B (32 x 32, pitch lb).
    // Every operand is loaded before the first product: loaded one step
    // ahead, each pair of products waited out a full LDS latency.
    __device__ void mac(const float* A, int la, const float* B, int lb, float s, int lane)
    {
        const int r = lane & 31, kk = lane >> 5;
        float av[16], bv[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            av[k] = A[r * la + 2 * k + kk];
            bv[k] = B[(2 * k + kk) * lb + r];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) c = __builtin_amdgcn_mfma_f32_32x32x2f32(s * av[k], bv[k], c, 0, 0, 0);
    }
    __device__ void load(const float* P, int lp, int lane)
    {
#pragma unroll
        for (int i = 0; i < 16; ++i) c[i] = P[row(lane, i) * lp + col(lane)];
    }
    __device__ void store(float* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 16; ++i) P[row(lane, i) * lp + col(lane)] = c[i];
    }
    // P -= this block (one rounding per entry)
    __device__ void sub_store(float* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 16; ++i) P[row(lane, i) * lp + col(lane)] -= c[i];
    }
    __device__ float absmax() const
    {
        float m = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; ++i) m = fmaxf(m, fabsf(c[i]));
        return m;
    }
};

template <>
struct Acc32<double> {
    mfma::doublex4 c[2][2];
    static __device__ int row(int lane, int b, int i) { return 16 * b + 4 * i + lane / 16; }
    static __device__ int col(int lane, int b) { return 16 * b + (lane & 15); }
    __device__ void zero()
    {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) c[a][b] = mfma::doublex4{0};
    }
    // operands loaded a half (K = 16) at a time ahead of its products
    __device__ void mac(const double* A, int la, const double* B, int lb, double s, int lane)
    {
        const int r = lane & 15, kk = lane >> 4;
#pragma unroll
        for (int h = 0; h < 32; h += 16) {
            double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = h + 4 * q + kk;
                a0[q] = A[r * la + k];
                a1[q] = A[(16 + r) * la + k];
                b0[q] = B[k * lb + r];
                b1[q] = B[k * lb + 16 + r];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(s * a0[q], b0[q], c[0][0], 0, 0, 0);
                c[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(s * a0[q], b1[q], c[0][1], 0, 0, 0);
                c[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(s * a1[q], b0[q], c[1][0], 0, 0, 0);
                c[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(s * a1[q], b1[q], c[1][1], 0, 0, 0);
            }
        }
    }
    __device__ void load(const double* P, int lp, int lane)
    {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) c[a][b][i] = P[row(lane, a, i) * lp + col(lane, b)];
    }
    __device__ void store(double* P, int lp, int lane) const
    {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) P[row(lane, a, i) * lp + col(lane, b)] = c[a][b][i];
    }
    __device__ void sub_store(double* P, int lp, int lane) const
    {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) P[row(lane, a, i) * lp + col(lane, b)] -= c[a][b][i];
    }
    __device__ double absmax() const
    {
        double m = 0.0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) m = fmax(m, fabs(c[a][b][i]));
        return m;
    }
};

// LU without interchanges of the 32 x 32 diagonal sub-block at (o, o) of S
// (pitch ls) by one wave: row o + r on lane r (lanes 32 .. 63 repeat row
// lane % 32 and store nothing), each pivot row broadcast with readlane (no
// barrier per column).  Returns the lane's largest |multiplier| (infinite
// for a zero pivot, left to the pivoting tiers).
template <typename T>
__device__ __forceinline__ T leaf_lu32(T* S, int ls, int o, int lane)
{
    T lmax = T(0);
    T a[SUBB];
    const int r = o + (lane & (SUBB - 1));
#pragma unroll
    for (int j = 0; j < SUBB; ++j) a[j] = S[r * ls + o + j];
#pragma unroll
    for (int c = 0; c < SUBB; ++c) {
        const T pv = read_lane(a[c], c);
        if (pv == T(0)) lmax = T(INFINITY);
        const T rp = pv != T(0) ? recip(pv) : T(1);
        const bool below = (lane & (SUBB - 1)) > c && lane < SUBB;
        const T l = below ? a[c] * rp : T(0);
        if (below) {
            a[c] = l;
            lmax = fmax(lmax, l < T(0) ? -l : l);
        }
        // the pivot row's values first, then the updates, sixteen columns at a
        // time: a readlane right after the FMA that wrote its source, or an
        // FMA right after the readlane it reads, costs wait states
        T pr[SUBB];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int j = c + 1 + h * 16; j < SUBB && j < c + 17 + h * 16; ++j) pr[j] = read_lane(a[j], c);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = c + 1 + h * 16; j < SUBB && j < c + 17 + h * 16; ++j) a[j] = fma(-l, pr[j], a[j]);
        }
    }
    if (lane < SUBB)
#pragma unroll
        for (int j = 0; j < SUBB; ++j) S[r * ls + o + j] = a[j];
    return lmax;
}

// One wave's 16 x 16 MFMA accumulator over K = 16 (v_mfma_f32_16x16x4_f32 /
// v_mfma_f64_16x16x4_f64: A[i][k] on lane i + 16 (k % 4), B[k][j] on lane
// j + 16 (k % 4); the result's row i, column j is on lane j + 16 (i / 4),
// register i % 4 in FP32, lane j + 16 (i % 4), register i / 4 in FP64), for
// the 16 x 16 blocks of the leaf inverses.
template <typename T>
struct Acc16;

template <>
struct Acc16<float> {
    typedef float floatx4 __attribute__((ext_vector_type(4)));
    floatx4 c;
    static __device__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
    __device__ void zero() { c = floatx4{0}; }
    __device__ void mac(const float* A, int la, const float* B, int lb, float s, int lane)
    {
        const int r = lane & 15, kk = lane >> 4;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(s * A[r * la + k0 + kk], B[(k0 + kk) * lb + r], c, 0, 0, 0);
    }
    __device__ void store(float* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) P[row(lane, i) * lp + (lane & 15)] = c[i];
    }
    __device__ void sub_store(float* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) P[row(lane, i) * lp + (lane & 15)] -= c[i];
    }
    __device__ float absmax() const
    {
        return fmaxf(fmaxf(fabsf(c[0]), fabsf(c[1])), fmaxf(fabsf(c[2]), fabsf(c[3])));
    }
};

template <>
struct Acc16<double> {
    mfma::doublex4 c;
    static __device__ int row(int lane, int i) { return 4 * i + (lane >> 4); }
    __device__ void zero() { c = mfma::doublex4{0}; }
    __device__ void mac(const double* A, int la, const double* B, int lb, double s, int lane)
    {
        const int r = lane & 15, kk = lane >> 4;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4)
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(s * A[r * la + k0 + kk], B[(k0 + kk) * lb + r], c, 0, 0, 0);
    }
    __device__ void store(double* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) P[row(lane, i) * lp + (lane & 15)] = c[i];
    }
    __device__ void sub_store(double* P, int lp, int lane) const
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) P[row(lane, i) * lp + (lane & 15)] -= c[i];
    }
    __device__ double absmax() const
    {
        return fmax(fmax(fabs(c[0]), fabs(c[1])), fmax(fabs(c[2]), fabs(c[3])));
    }
};

template <typename T>
__global__ __launch_bounds__(256) void diag_block_kernel(T* A, int64_t lda, int nb, int64_t k0, int32_t* ipiv,
                                                         unsigned int* lmax_bits, T* invL, T* invU, int64_t zb)
{
    A = zoff(A, zb);
    ipiv = zoff(ipiv, zb);
    lmax_bits = zoff(lmax_bits, zb);
    invL = zoff(invL, zb);
    invU = zoff(invU, zb);

    constexpr int LS = NB + 1, LB = SUBB + 1, NBB = NB / SUBB;
    // FP32 keeps every leaf inverse and the assembled inverse blocks in LDS
    // (IL_k, IU_k, then the six off-diagonal blocks of each inverse); FP64's
    // block fills most of LDS, so its leaf inverses go through LDS one at a
    // time and the assembly reads and writes invL and invU in global memory.
    constexpr bool KEEP = sizeof(T) == 4;
    constexpr int NIL = KEEP ? NBB : 1, NX = KEEP ? 6 : 1, XSZ = KEEP ? SUBB * LB : 1;
    __shared__ T S[NB * LS];
    __shared__ T ILs[NIL][SUBB * LB], IUs[NIL][SUBB * LB];
    __shared__ T XLs[NX][XSZ], YUs[NX][XSZ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    load_diag_block<T>(A, lda, nb, S, tid);
    for (int c = tid; c < nb; c += 256) ipiv[k0 + c] = (int32_t)(k0 + c);
    __syncthreads();
    T lmax = T(0);
    // The inverses' off-diagonal blocks: X = inv(L) block (i, j), j < i, and
    // Y = inv(U) block (j, i), have index q(i, j) = i (i - 1) / 2 + j (FP32:
    // XLs / YUs; FP64: in place in invL / invU); the leaf inverses IL_i, IU_i
    // are their diagonal blocks.
    auto qi = [](int i, int j) { return i * (i - 1) / 2 + j; };
    auto ilp = [&](int i) -> const T* { return KEEP ? ILs[i] : invL + i * SUBB * NB + i * SUBB; };
    auto iup = [&](int i) -> const T* { return KEEP ? IUs[i] : invU + i * SUBB * NB + i * SUBB; };
    constexpr int LP = KEEP ? LB : NB;   // pitch of those blocks
    auto xp = [&](int i, int j) -> T* { return KEEP ? XLs[qi(i, j)] : invL + i * SUBB * NB + j * SUBB; };
    auto yp = [&](int j, int i) -> T* { return KEEP ? YUs[qi(i, j)] : invU + j * SUBB * NB + i * SUBB; };
    // A_ib -= L_ik U_kb at step k (the product summed from zero, then
    // subtracted: one rounding at the scale of A_ib)
    auto update = [&](int k, int ib, int jb) {
        Acc32<T> acc;
        acc.zero();
        acc.mac(S + ib * SUBB * LS + k * SUBB, LS, S + k * SUBB * LS + jb * SUBB, LS, T(1), lane);
        acc.sub_store(S + ib * SUBB * LS + jb * SUBB, LS, lane);
    };
    // Block row i of inv(L) and block column i of inv(U), task t < i:
    //   X_it = -IL_i sum_{m = t}^{i - 1} L_im X_mt,
    // task t >= i (j = t - i):
    //   Y_ji = -(sum_{m = j}^{i - 1} Y_jm U_mi) IU_i,
    // with X_mm = IL_m, Y_mm = IU_m: the sum into the output's place, then the
    // outer product in place (the wave reads all its operands before it stores).
    // (a wave's own earlier stores are read back by other lanes: fenced)
    auto wave_fence = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    auto inv_sum = [&](int i, int t) {
        wave_fence();
        Acc32<T> acc;
        acc.zero();
        const bool lo = t < i;
        const int j = lo ? t : t - i;
        for (int m = j; m < i; ++m) {
            if (lo) acc.mac(S + i * SUBB * LS + m * SUBB, LS, m == j ? ilp(j) : xp(m, j), LP, T(1), lane);
            else acc.mac(m == j ? iup(j) : yp(j, m), LP, S + m * SUBB * LS + i * SUBB, LS, T(1), lane);
        }
        acc.store(lo ? xp(i, j) : yp(j, i), LP, lane);
    };
    auto inv_outer = [&](int i, int t) {
        wave_fence();
        Acc32<T> acc;
        acc.zero();
        const bool lo = t < i;
        const int j = lo ? t : t - i;
        T* out = lo ? xp(i, j) : yp(j, i);
        if (lo) acc.mac(ilp(i), LP, out, LP, T(-1), lane);
        else acc.mac(out, LP, iup(i), LP, T(-1), lane);
        acc.store(out, LP, lane);
    };
    // The factors -> A, rows [r0, r1) x columns [c0, c1) of the block, by
    // thread t0 of nt (16-byte stores when the block is whole and aligned; the
    // LDS copy's rows are not 16-byte aligned: scalar reads)
    using VT = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    const bool vec_store = nb == NB && lda % VW == 0 && ((uintptr_t)A % 16) == 0;
    auto store_factors = [&](int r0, int r1, int c0, int c1, int t0, int nt) {
        const int rv = (c1 - c0) / VW;
        for (int idx = t0; idx < (r1 - r0) * rv; idx += nt) {
            const int rr = r0 + idx / rv, c = c0 + (idx % rv) * VW;
            VT v;
            T* vp = reinterpret_cast<T*>(&v);
#pragma unroll
            for (int q = 0; q < VW; ++q) vp[q] = S[rr * LS + c + q];
            *reinterpret_cast<VT*>(A + rr * lda + c) = v;
        }
    };
    // Each step kb opens with wave 0 updating diagonal block kb by step kb - 1
    // and factoring its leaf, while waves 1 .. 3 do the rest of step kb - 1:
    // its other trailing updates and block row / column kb - 1 of the
    // inverses (at the last leaf also the sums of row / column NBB - 1 that
    // need no IL_(NBB-1), and the finished factor blocks).  Then all four waves
    // invert the leaf and form the panels of step kb.
    static_assert(NBB == 4, "the last step's work lists are laid out for four leaves");
    for (int kb = 0;; ++kb) {
        const int o = kb * SUBB;
        T* IL = ILs[KEEP ? kb : 0];
        T* IU = IUs[KEEP ? kb : 0];
        if (wave == 0) {
            if (kb > 0) update(kb - 1, kb, kb);
            lmax = fmax(lmax, leaf_lu32<T>(S, LS, o, lane));
        } else if (kb > 0 && kb < NBB - 1) {
            const int k = kb - 1, rest = NBB - kb;
            const int nu = rest * rest - 1;
#pragma unroll 1
            for (int q = wave - 1; q < nu + 2 * k; q += 3) {
                if (q < nu) {
                    update(k, kb + (q + 1) / rest, kb + (q + 1) % rest);
                } else {
                    inv_sum(k, q - nu);
                    inv_outer(k, q - nu);
                }
            }
        } else if (kb == NBB - 1) {
            // X_20, X_21, Y_02, Y_12 and the sums of X_3j, Y_j3 (each after the
            // row-2 block it reads, on the same wave; (whole task?, i, t) per
            // entry, 7 / 8 / 7 products); then the factors but block (3, 3)
            constexpr signed char work[3][4][3] = {{{1, 2, 0}, {0, 3, 0}, {0, 3, 2}, {-1, 0, 0}},
                                                   {{1, 2, 1}, {0, 3, 1}, {1, 2, 3}, {0, 3, 4}},
                                                   {{1, 2, 2}, {0, 3, 3}, {0, 3, 5}, {-1, 0, 0}}};
#pragma unroll 1
            for (int e = 0; e < 4; ++e) {
                const int whole = work[wave - 1][e][0], i = work[wave - 1][e][1], t = work[wave - 1][e][2];
                if (whole < 0) break;
                inv_sum(i, t);
                if (whole) inv_outer(i, t);
            }
            if (vec_store) {
                store_factors(0, o, 0, NB, tid - 64, 192);
                store_factors(o, NB, 0, o, tid - 64, 192);
            }
        }
        __syncthreads();
        const T* F = S + o * LS + o;
        {
            // The leaf's triangles in 16 x 16 halves: wave 0 inverts L11, wave 1
            // L22, wave 2 U11, wave 3 U22 (one column per lane, by
            // substitution: chains of 120 dependent steps instead of 496);
            // waves 0 and 2 also zero IL12 and IU21.
            constexpr int H = SUBB / 2;
            const int h = (wave & 1) * H;
            const T* Fh = F + h * LS + h;
            if (lane < H) {
                // the triangle's entries are loaded into registers before the
                // substitution (read as it went, each row waited out an LDS
                // latency); f holds row i's entries m < i (L) or row
                // H - 1 - i's entries H - 1 - m, m < i, then its diagonal (U)
                constexpr int NT = H * (H - 1) / 2;
                const int j = lane;
                T f[NT + H];
                if (wave < 2) {
#pragma unroll
                    for (int i = 1, e = 0; i < H; ++i)
#pragma unroll
                        for (int m = 0; m < i; ++m, ++e) f[e] = Fh[i * LS + m];
                    T x[H];
#pragma unroll
                    for (int i = 0, e = 0; i < H; ++i) {
                        T acc = T(0);
#pragma unroll
                        for (int m = 0; m < i; ++m, ++e) acc = fma(f[e], x[m], acc);
                        x[i] = i < j ? T(0) : (i == j ? T(1) : -acc);
                    }
#pragma unroll
                    for (int i = 0; i < H; ++i) IL[(h + i) * LB + h + j] = x[i];
                } else {
#pragma unroll
                    for (int q = 1, e = 0; q < H; ++q)
#pragma unroll
                        for (int m = 0; m < q; ++m, ++e) f[e] = Fh[(H - 1 - q) * LS + H - 1 - m];
#pragma unroll
                    for (int i = 0; i < H; ++i) f[NT + i] = Fh[i * LS + i];
                    // y[q] holds row H - 1 - q
                    T y[H];
#pragma unroll
                    for (int q = 0, e = 0; q < H; ++q) {
                        const int i = H - 1 - q;
                        T acc = T(0);
#pragma unroll
                        for (int m = 0; m < q; ++m, ++e) acc = fma(f[e], y[m], acc);
                        const T rd = recip(f[NT + i]);
                        y[q] = i > j ? T(0) : (i == j ? rd : -acc * rd);
                    }
#pragma unroll
                    for (int q = 0; q < H; ++q) IU[(h + H - 1 - q) * LB + h + j] = y[q];
                }
            } else if (lane < 2 * H && (wave & 1) == 0) {
                const int j = lane - H;
#pragma unroll
                for (int i = 0; i < H; ++i) {
                    if (wave == 0) IL[i * LB + H + j] = T(0);
                    else IU[(H + i) * LB + j] = T(0);
                }
            }
            __syncthreads();
            // IL21 = -IL22 (L21 IL11) on wave 0, IU12 = -(IU11 U12) IU22 on
            // wave 1: the inner product into the output's place, then the outer
            // one in place (a wave reads all its operands before it stores)
            if (wave < 2) {
                Acc16<T> p;
                p.zero();
                if (wave == 0) p.mac(F + H * LS, LS, IL, LB, T(1), lane);
                else p.mac(IU, LB, F + H, LS, T(1), lane);
                T* out = wave == 0 ? IL + H * LB : IU + H;
                p.store(out, LB, lane);
                p.zero();
                if (wave == 0) p.mac(IL + H * LB + H, LB, out, LB, T(-1), lane);
                else p.mac(out, LB, IU + H * LB + H, LB, T(-1), lane);
                p.store(out, LB, lane);
            }
        }
        __syncthreads();
        if (!KEEP) {
            // FP64: the leaf inverses are the diagonal blocks of invL and invU
            for (int e = tid; e < SUBB * SUBB; e += 256) {
                const int i = e / SUBB, j = e % SUBB;
                invL[(o + i) * NB + o + j] = IL[i * LB + j];
                invU[(o + i) * NB + o + j] = IU[i * LB + j];
            }
        }
        if (kb == NBB - 1) break;
        const int rest = NBB - 1 - kb;
        // panels L_bk = A_bk IU (task p even, b = kb + 1 + p / 2) and
        // U_kb = IL A_kb (p odd): those of block kb + 1 first, on waves 0 and 1
        for (int p = wave; p < 2 * rest; p += 4) {
            const int b = kb + 1 + p / 2;
            Acc32<T> acc;
            acc.zero();
            if ((p & 1) == 0) {
                T* blk = S + b * SUBB * LS + o;
                acc.mac(blk, LS, IU, LB, T(1), lane);
                acc.store(blk, LS, lane);
                lmax = fmax(lmax, acc.absmax());
            } else {
                T* blk = S + o * LS + b * SUBB;
                acc.mac(IL, LB, blk, LS, T(1), lane);
                acc.store(blk, LS, lane);
            }
        }
        __syncthreads();
    }
    // the products by IL_3 and IU_3 of row / column 3 of the inverses (FP64
    // reads IL_3, IU_3 from the copies in invL, invU); the factors of block
    // (3, 3) (all of them when the block is partial)
    if (!KEEP) __syncthreads();
#pragma unroll 1
    for (int t = wave; t < 2 * (NBB - 1); t += 4) inv_outer(NBB - 1, t);
    if (vec_store) {
        store_factors((NBB - 1) * SUBB, NB, (NBB - 1) * SUBB, NB, tid, 256);
    } else if (nb == NB) {
        for (int r = tid / NB; r < NB; r += 256 / NB) A[r * lda + tid % NB] = S[r * LS + tid % NB];
    } else {
        for (int r = 0; r < nb; ++r)
            for (int c = tid; c < nb; c += 256) A[r * lda + c] = S[r * LS + c];
    }
    __syncthreads();
    // largest multiplier of the block -> *lmax_bits
    const T wl = wave_max(lmax);
    if (lane == 0) atomicMax(lmax_bits, __float_as_uint((float)wl));
    // the inverses -> invL, invU, block by block (FP32 from LDS; FP64's
    // nonzero blocks are already in place, only its zero blocks are written)
    for (int blk = 0; blk < NBB * NBB; ++blk) {
        const int bi = blk / NBB, bj = blk % NBB;
        const T* srcL = nullptr;
        const T* srcU = nullptr;
        if (KEEP) {
            srcL = bj > bi ? nullptr : (bj == bi ? ILs[bi] : XLs[qi(bi, bj)]);
            srcU = bj < bi ? nullptr : (bj == bi ? IUs[bi] : YUs[qi(bj, bi)]);
        } else if (bj <= bi && bj >= bi) {
            continue;   // FP64 diagonal blocks: both inverses already written
        }
        T* dL = invL + bi * SUBB * NB + bj * SUBB;
        T* dU = invU + bi * SUBB * NB + bj * SUBB;
        // 16-byte stores (the blocks start 16-byte aligned in the aux)
        for (int e = tid; e < SUBB * SUBB / VW; e += 256) {
            const int ri = e / (SUBB / VW), cj = (e % (SUBB / VW)) * VW;
            VT vl, vu;
            T* pl = reinterpret_cast<T*>(&vl);
            T* pu = reinterpret_cast<T*>(&vu);
#pragma unroll
            for (int q = 0; q < VW; ++q) {
                pl[q] = srcL ? srcL[ri * LB + cj + q] : T(0);
                pu[q] = srcU ? srcU[ri * LB + cj + q] : T(0);
            }
            if (KEEP || bj > bi) *reinterpret_cast<VT*>(dL + ri * NB + cj) = vl;
            if (KEEP || bj < bi) *reinterpret_cast<VT*>(dU + ri * NB + cj) = vu;
        }
    }
}

// info = the first zero pivot: keeps an earlier one.
__global__ void fold_info_kernel(const int32_t* blk, int32_t* info)
{
    if (*blk != 0) atomicCAS(info, 0, *blk);
}

// ---------------------------------------------------------------- row interchanges
// Applies ipiv[k .. k+nb) in order to columns [c0, c1) of A (one lane per column).
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(T* A, int64_t lda, int64_t k, int nb, const int32_t* ipiv,
                                                    int64_t c0, int64_t c1)
{
    const int64_t col = c0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= c1) return;
    for (int r = 0; r < nb; ++r) {
        const int64_t p = ipiv[k + r];
        if (p != k + r) {
            T t = A[(k + r) * lda + col];
            A[(k + r) * lda + col] = A[p * lda + col];
            A[p * lda + col] = t;
        }
    }
}

// ---------------------------------------------------------------- general pivoting path
// Column c of the panel [k, k+nb): pivot over rows [c, n), swap within the
// panel columns, scale the multipliers.  One workgroup of 1024 lanes.
template <typename T>
__global__ __launch_bounds__(1024) void pivot_col_kernel(T* A, int64_t lda, int64_t n, int64_t k, int nb, int64_t c,
                                                         int32_t* ipiv, int32_t* info)
{
    __shared__ T red_val[16];
    __shared__ int64_t red_idx[16];
    __shared__ int64_t piv_s;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    T v = T(-1);
    int64_t vi = c;
    for (int64_t r = c + tid; r < n; r += 1024) {
        T a = (T)fabs((double)A[r * lda + c]);
        if (a > v) {
            v = a;
            vi = r;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off /= 2) {
        T ov = __shfl_xor(v, off);
        int64_t oi = __shfl_xor(vi, off);
        if (ov > v || (ov == v && oi < vi)) {
            v = ov;
            vi = oi;
        }
    }
    if (lane == 0) {
        red_val[wave] = v;
        red_idx[wave] = vi;
    }
    __syncthreads();
    if (tid == 0) {
        T bv = red_val[0];
        int64_t bi = red_idx[0];
        for (int w = 1; w < 16; ++w)
            if (red_val[w] > bv || (red_val[w] == bv && red_idx[w] < bi)) {
                bv = red_val[w];
                bi = red_idx[w];
            }
        piv_s = bi;
        ipiv[c] = (int32_t)bi;
        if (A[bi * lda + c] == T(0)) atomicCAS(info, 0, (int32_t)(c + 1));
    }
    __syncthreads();
    const int64_t p = piv_s;
    if (p != c && tid < nb) {
        T t = A[c * lda + k + tid];
        A[c * lda + k + tid] = A[p * lda + k + tid];
        A[p * lda + k + tid] = t;
    }
    __syncthreads();
    const T pv = A[c * lda + c];
    if (pv != T(0)) {
        const T rp = T(1) / pv;
        for (int64_t r = c + 1 + tid; r < n; r += 1024) A[r * lda + c] *= rp;
    }
}

// Rank-1 update of the panel columns (c, k+nb) for rows (c, n).
template <typename T>
__global__ __launch_bounds__(256) void panel_update_kernel(T* A, int64_t lda, int64_t n, int64_t k, int nb, int64_t c)
{
    const int64_t r0 = c + 1 + (int64_t)blockIdx.x * 64;
    const int w = (int)(k + nb - c - 1);
    if (w <= 0) return;
    for (int idx = threadIdx.x; idx < 64 * w; idx += 256) {
        const int64_t r = r0 + idx / w;
        const int64_t j = c + 1 + idx % w;
        if (r < n) A[r * lda + j] -= A[r * lda + c] * A[c * lda + j];
    }
}

// ---------------------------------------------------------------- chained block substitution
// Forward (LOWER: L z = y, unit diagonal) or backward (U x = z) substitution
// by NB-row blocks in one launch, for NR-padded right-hand sides stored
// row-major with ld nrhs <= NR.  Block row p needs every block solved before
// it in the sweep, written with its two nearest blocks q1 = p -/+ 1 and
// q2 = p -/+ 2 apart:
//   c_p = inv(T_pp) (b_p - sum_{j != q1, q2} T_pj x_j)   (j < p forward, j > p backward),
//   x_p = c_p - M2_p x_q2 - M_p x_q1,  M_p = inv(T_pp) T_pq1, M2_p = inv(T_pp) T_pq2
// (inverse and links kept in aux), so c_p and its far sums are formed while
// the two blocks before it are still being solved, and the step from x_q1 to
// x_p is one NB x NB product.
// Workgroups take tickets in dispatch order (ticket t solves block t forward,
// nblk - 1 - t backward), so a workgroup waits only on workgroups dispatched
// before it.  The output x is filled with a signalling-NaN sentinel before
// the launch (arithmetic never produces one); a workgroup publishes x_p by
// storing it, waiting for its stores and issuing one agent-scope release
// (the L2 write-back that makes them visible to the other XCDs), and a
// reader polls the entries of x_j it needs with coherent loads until none is
// the sentinel: the data is the flag, so a block step costs one round trip
// through memory.  A wait is bounded (CH_SPIN polls): a workgroup that gives
// up sets sync[1] and writes quiet NaN for its block, so every workgroup ends.
// sync = [ticket, error], zeroed before the launch.
constexpr int CH_SPIN = 1 << 22;
// NB x NB blocks per diagonal block in the factor's aux (FactorAux): FP32
// factors keep the two-step links ML2, MU2 as well (trsm_chain_kernel)
template <typename T>
constexpr int aux_blocks()
{
    return sizeof(T) == 4 ? 6 : 4;
}

__global__ void zero_u32_kernel(unsigned* p, int64_t n, int64_t zb)
{
    p = zoff(p, zb);

    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The chain's "not yet written" value: a signalling NaN of each width.
template <typename T>
struct ChainBits;
template <>
struct ChainBits<float> {
    using U = unsigned int;
    static constexpr U sentinel = 0x7FA5A5A5u;
};
template <>
struct ChainBits<double> {
    using U = unsigned long long;
    static constexpr U sentinel = 0x7FF4A5A5A5A5A5A5ull;
};

template <typename T>
__global__ void fill_sentinel_kernel(T* x, int64_t n, int64_t zb)
{
    x = zoff(x, zb);

    using U = typename ChainBits<T>::U;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) reinterpret_cast<U*>(x)[i] = ChainBits<T>::sentinel;
}

// Bounded wait for one entry of x written by another workgroup.
template <typename T>
__device__ __forceinline__ T chain_read(const T* p, bool& gave_up)
{
    using U = typename ChainBits<T>::U;
    const U* q = reinterpret_cast<const U*>(p);
    U v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = gave_up ? CH_SPIN : 0;   // a thread waits out the limit at most once
    while (v == ChainBits<T>::sentinel) {
        if (++spins > CH_SPIN) {
            gave_up = true;
            return T(NAN);
        }
        __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    T out;
    __builtin_memcpy(&out, &v, sizeof(T));
    return out;
}

// Sum over the TPR consecutive lanes that share a row (TPR = 2 or 4), by DPP.
template <int TPR, typename T>
__device__ __forceinline__ T row_sum(T v)
{
    v += dpp_mov<0xB1>(v);
    if (TPR == 4) v += dpp_mov<0x4E>(v);
    return v;
}

// TPR lanes per row of the block (NB x TPR lanes per workgroup), each
// covering NB / TPR columns of every block product.
template <typename T, bool LOWER, int NR, int TPR>
__global__ __launch_bounds__(NB* TPR) void trsm_chain_kernel(const T* LU, int64_t lda, int64_t n, const T* invs,
                                                             const T* b, T* x, int nrhs, unsigned* sync, int64_t zb)
{
    LU = zoff(LU, zb);
    invs = zoff(invs, zb);
    b = zoff(b, zb);
    x = zoff(x, zb);
    sync = zoff(sync, zb);

    constexpr int H = NB / TPR;
    constexpr int NT = NB * TPR;
    constexpr int V = Vec16<T>::n;
    constexpr int TP = NB + (sizeof(T) == 4 ? 4 : 2);  // LDS row pitch: rows 16-byte aligned, 4 banks apart
    using VT = typename Vec16<T>::type;
    __shared__ __attribute__((aligned(16))) T Ts[NB * TP];
    __shared__ __attribute__((aligned(16))) T xs[NB * NR];
    __shared__ int tk;
    const int tid = threadIdx.x, r = tid / TPR, hf = tid % TPR;
    const int64_t nblk = (n + NB - 1) / NB;
    if (tid == 0) tk = (int)atomicAdd(&sync[0], 1u);
    __syncthreads();
    const int64_t p = LOWER ? tk : nblk - 1 - tk;
    const int64_t r0 = p * NB;
    const int nbp = (int)min((int64_t)NB, n - r0);
    bool gave_up = false;
    T part[NR], bq[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        part[q] = T(0);
        // this row's right-hand side, read now: it is not produced by the chain
        bq[q] = (hf == 0 && r < nbp && q < nrhs) ? b[(r0 + r) * nrhs + q] : T(0);
    }
    const int64_t nsteps = LOWER ? p : nblk - 1 - p;
    // this thread's part of the row of T_pj: all H entries loaded at once
    // (issued a step ahead, so the loads overlap the previous block's products)
    T trow[H];
    auto load_row = [&](int64_t j) {
        const int64_t c0 = j * NB;
        const int nbj = (int)min((int64_t)NB, n - c0);
        const T* src = LU + (r0 + min(r, nbp - 1)) * lda + c0 + hf * H;
        const bool full = r < nbp && nbj == NB && ((uintptr_t)src % 16) == 0;
        if (full) {
#pragma unroll
            for (int k = 0; k < H; k += V) {
                const VT u = *reinterpret_cast<const VT*>(src + k);
                const T* up = reinterpret_cast<const T*>(&u);
#pragma unroll
                for (int q = 0; q < V; ++q) trow[k + q] = up[q];
            }
        } else {
#pragma unroll
            for (int k = 0; k < H; ++k) trow[k] = (r < nbp && hf * H + k < nbj) ? src[k] : T(0);
        }
    };
    // the diagonal block's inverse row and the link row M_p (always NB x NB,
    // zero or identity beyond the block; rows of acc beyond it are zero)
    const T* blk = invs + p * aux_blocks<T>() * NB * NB + (int64_t)r * NB + hf * H;
    const T* inv = blk + (LOWER ? 0 : NB * NB);
    const T* lnk = blk + (LOWER ? 2 : 3) * NB * NB;
    const T* lnk2 = blk + (LOWER ? 4 : 5) * NB * NB;
    auto load_nb = [&](const T* src) {
#pragma unroll
        for (int k = 0; k < H; k += V) {
            const VT u = *reinterpret_cast<const VT*>(src + k);
            const T* up = reinterpret_cast<const T*>(&u);
#pragma unroll
            for (int q = 0; q < V; ++q) trow[k + q] = up[q];
        }
    };
    auto stage_row = [&]() {
        VT* dst = reinterpret_cast<VT*>(&Ts[r * TP + hf * H]);
#pragma unroll
        for (int k = 0; k < H; k += V) {
            VT u;
            T* up = reinterpret_cast<T*>(&u);
#pragma unroll
            for (int q = 0; q < V; ++q) up[q] = trow[k + q];
            dst[k / V] = u;
        }
    };
    // acc[q] += (this thread's columns of the staged row) . xs[:, q]
    auto row_product = [&](T* acc) {
        const T* tr = &Ts[r * TP + hf * H];
#pragma unroll 1
        for (int k0 = 0; k0 < H; k0 += V) {
            const VT u = *reinterpret_cast<const VT*>(tr + k0);
            const T* up = reinterpret_cast<const T*>(&u);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const T* xk = &xs[(hf * H + k0 + k) * NR];
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = fma(up[k], xk[q], acc[q]);
            }
        }
    };
    // x_j of block j of the sweep into xs (waiting for each entry)
    auto read_x = [&](int64_t j) {
        const int64_t c0 = j * NB;
        const int nbj = (int)min((int64_t)NB, n - c0);
        for (int idx = tid; idx < NB * NR; idx += NT) {
            const int c = idx / NR, q = idx % NR;
            xs[idx] = (c < nbj && q < nrhs) ? chain_read(&x[(c0 + c) * nrhs + q], gave_up) : T(0);
        }
    };
    // every block of the sweep before this one but the nearest (FP32: but the
    // two nearest, which gives the far sums and c_p two block steps of the
    // chain to finish in instead of one; FP64 has no registers left for the
    // second link's row): acc -= T_pj x_j
    constexpr bool TWO = sizeof(T) == 4;
    const int64_t far = TWO ? (nsteps > 1 ? nsteps - 2 : 0) : (nsteps > 0 ? nsteps - 1 : 0);
    if (far > 0) load_row(LOWER ? 0 : nblk - 1);
    else load_nb(inv);
#pragma unroll 1
    for (int64_t s = 0; s < far; ++s) {
        stage_row();
        read_x(LOWER ? s : nblk - 1 - s);
        __syncthreads();
        if (s + 1 < far) load_row(LOWER ? s + 1 : nblk - 2 - s);
        else load_nb(inv);
        row_product(part);
        __syncthreads();
    }
    // c_p = inv(T_pp) (b_p - acc), formed while the nearest block is still
    // being solved by the workgroup before this one
#pragma unroll
    for (int q = 0; q < NR; ++q) part[q] = row_sum<TPR>(part[q]);  // the TPR parts of row r
    if (hf == 0)
#pragma unroll
        for (int q = 0; q < NR; ++q)
            xs[r * NR + q] = (r < nbp && q < nrhs) ? bq[q] - part[q] : T(0);
    stage_row();
    __syncthreads();
    // the links' rows, loaded while c_p is formed: M2_p into trow and M_p
    // into mrow when the second nearest block joins through its link
    const bool two = TWO && nsteps > 1;
    T mrow[TWO ? H : 1];
    if (two) {
        load_nb(lnk2);
#pragma unroll
        for (int k = 0; k < (TWO ? H : 1); k += V) {
            const VT u = *reinterpret_cast<const VT*>(lnk + k);
            const T* up = reinterpret_cast<const T*>(&u);
#pragma unroll
            for (int q = 0; q < V; ++q) mrow[k + q] = up[q];
        }
    } else if (nsteps > 0) {
        load_nb(lnk);
    }
    T y[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) y[q] = T(0);
    row_product(y);
#pragma unroll
    for (int q = 0; q < NR; ++q) y[q] = row_sum<TPR>(y[q]);
    // y = c_p - M2_p x_second - M_p x_nearest
    auto link_step = [&](int64_t j) {
        __syncthreads();
        stage_row();
        read_x(j);
        __syncthreads();
        T z[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) z[q] = T(0);
        row_product(z);
#pragma unroll
        for (int q = 0; q < NR; ++q) y[q] -= row_sum<TPR>(z[q]);
    };
    if (two) {
        link_step(LOWER ? p - 2 : p + 2);
#pragma unroll
        for (int k = 0; k < (TWO ? H : 1); ++k) trow[k] = mrow[k];
    }
    // the only work after the nearest block arrives: one NB x NB product
    if (nsteps > 0) link_step(LOWER ? p - 1 : p + 1);
    const int bad = __syncthreads_or(gave_up);
    if (tid == 0 && bad) atomicOr(&sync[1], 1u);
    if (hf == 0 && r < nbp)
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nrhs) x[(r0 + r) * nrhs + q] = bad ? T(NAN) : y[q];
    // every wave's stores complete (reach this XCD's L2), then one release:
    // its L2 write-back publishes all of them to the other XCDs (cheaper than
    // agent-scope atomic stores entry by entry, measured)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// dst[i][:] = scale_c * src[perm[i]][:] with a type conversion (gather of the
// right-hand side by the factor's row permutation).
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void gather_rows_kernel(const TS* src, int64_t lds, TD* dst, int64_t n, int nrhs,
                                                          const int32_t* perm, const double* inv_scale, int64_t zb)
{
    src = zoff(src, zb);
    dst = zoff(dst, zb);
    perm = zoff(perm, zb);
    inv_scale = zoff(inv_scale, zb);

    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    const int64_t s = perm ? perm[i] : i;
    double v = (double)src[s * lds + q];
    if (inv_scale) v *= inv_scale[q];
    dst[idx] = (TD)v;
}

// Per-point 3x3 frame R (row-major, lab = R body) for refinement with factors
// of the body-frame matrix: the residual goes to the body frame (R^T r)
// before the FP32 solve and the correction back to the lab frame (R d).
struct Frame3 {
    double r[9];
};

// dst[i][q] = inv_scale[q] * (R^T src)[perm[i]][q] (perm == nullptr: the
// identity); rows are point-major (row 3p + a is component a of point p).
__global__ __launch_bounds__(256) void gather_rot_kernel(const double* src, float* dst, int64_t n,
                                                         int nrhs, const int32_t* perm, const double* inv_scale,
                                                         Frame3 R)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    const int64_t s = perm ? perm[i] : i;
    const int64_t p3 = s - s % 3;
    const int a = (int)(s % 3);
    const double v = fma(R.r[a], src[p3 * nrhs + q],
                         fma(R.r[3 + a], src[(p3 + 1) * nrhs + q], R.r[6 + a] * src[(p3 + 2) * nrhs + q]));
    dst[idx] = (float)(v * inv_scale[q]);
}

// x[3p + a][q] += scale[q] * (R d)[3p + a][q]; x and d have ld nrhs.
__global__ __launch_bounds__(256) void axpy_rot_kernel(double* x, const float* d, int64_t n, int nrhs,
                                                       const double* scale, Frame3 R)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    const int64_t p3 = i - i % 3;
    const int a = (int)(i % 3);
    const double v = fma(R.r[3 * a], (double)d[p3 * nrhs + q],
                         fma(R.r[3 * a + 1], (double)d[(p3 + 1) * nrhs + q], R.r[3 * a + 2] * (double)d[(p3 + 2) * nrhs + q]));
    x[idx] += scale[q] * v;
}

template <typename TS, typename TD>
__global__ __launch_bounds__(256) void copy_rows_kernel(const TS* src, TD* dst, int64_t ldd, int64_t n, int nrhs, int64_t zb)
{
    src = zoff(src, zb);
    dst = zoff(dst, zb);

    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    dst[i * ldd + q] = (TD)src[idx];
}

// ---------------------------------------------------------------- refinement
// r = b - A x (FP64), 8 rows per workgroup: each x entry is loaded once for
// the 8 rows, the four waves take interleaved quarters of the columns (lanes
// striding along a row, so A is read in 256-lane coalesced runs), and the
// partial sums are combined in a fixed order (wave 0 .. 3) through LDS.  At
// n = 5,400 that is 675 workgroups, enough to fill the chip.  x and r are
// n x nrhs (ld nrhs), b has ld ldb.
constexpr int RES_ROWS = 8;
// PAIR: two adjacent columns per lane and step, read as one 16-byte load per
// row (A 16-byte aligned, lda even), which doubles the bytes each load keeps
// in flight.  NQ = nrhs (1 .. 8) sizes the accumulators: with 6 right-hand
// sides the kernel stays under 256 VGPRs (two waves per SIMD, twice the
// loads in flight of the 8-column form, which would spill under that
// bound).  Either variant sums in a fixed order for a given shape.
template <bool PAIR, int NQ>
__global__ __launch_bounds__(256, NQ < 8 ? 2 : 1) void residual_kernel(const double* A, int64_t lda, int64_t n, const double* b,
                                                       int64_t ldb, const double* x, double* r, int64_t zb)
{
    A = zoff(A, zb);
    b = zoff(b, zb);
    x = zoff(x, zb);
    r = zoff(r, zb);

    __shared__ double part[4][RES_ROWS * NQ];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int64_t row0 = (int64_t)blockIdx.x * RES_ROWS;
    double acc[RES_ROWS][NQ];
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[i][q] = 0.0;
    int64_t c = PAIR ? 2 * (int64_t)threadIdx.x : (int64_t)threadIdx.x;
    if (PAIR) {
        for (; c + 1 < n; c += 512) {
            double x0[NQ], x1[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                x0[q] = x[c * NQ + q];
                x1[q] = x[(c + 1) * NQ + q];
            }
#pragma unroll
            for (int i = 0; i < RES_ROWS; ++i) {
                const double2 a = row0 + i < n ? *reinterpret_cast<const double2*>(A + (row0 + i) * lda + c)
                                               : make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < NQ; ++q) acc[i][q] += a.x * x0[q] + a.y * x1[q];
            }
        }
    }
    for (; c < n; c += PAIR ? 512 : 256) {   // PAIR: at most the odd last column
        double xv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) xv[q] = x[c * NQ + q];
#pragma unroll
        for (int i = 0; i < RES_ROWS; ++i) {
            const double a = row0 + i < n ? A[(row0 + i) * lda + c] : 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[i][q] += a * xv[q];
        }
    }
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = acc[i][q];
#pragma unroll
            for (int off = 32; off > 0; off /= 2) s += __shfl_xor(s, off);
            acc[i][q] = s;
        }
    if (lane < RES_ROWS * NQ) {
        double s = 0.0;
#pragma unroll
        for (int ii = 0; ii < RES_ROWS; ++ii)
#pragma unroll
            for (int qq = 0; qq < NQ; ++qq)
                if (ii * NQ + qq == lane) s = acc[ii][qq];
        part[wave][lane] = s;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < RES_ROWS * NQ) {
        const int i = t / NQ, q = t % NQ;
        const double s = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
        if (row0 + i < n) r[(row0 + i) * NQ + q] = b[(row0 + i) * ldb + q] - s;
    }
}

// r = b - A x through residual_kernel<PAIR, nrhs> (x and r with ld nrhs)
template <bool PAIR, int NQ = 1>
void residual_launch(int nrhs, dim3 grid, hipStream_t st, const double* A, int64_t lda, int64_t n, const double* b,
                     int64_t ldb, const double* x, double* r, int64_t zb)
{
    if constexpr (NQ < 8) {
        if (nrhs != NQ) return residual_launch<PAIR, NQ + 1>(nrhs, grid, st, A, lda, n, b, ldb, x, r, zb);
    }
    hipLaunchKernelGGL((residual_kernel<PAIR, NQ>), grid, dim3(256), 0, st, A, lda, n, b, ldb, x, r, zb);
}

// y = X t for an explicit FP32 inverse X (n x n, ldx) and nrhs <= 8 FP32
// columns t, y (ld nrhs): the correction of a refinement step with an
// explicit inverse.  One workgroup per RES_ROWS rows, columns split over the
// lanes, the four waves' partial sums combined in a fixed order.  Bound by
// reading X once (4 n^2 bytes).
__global__ __launch_bounds__(256) void inv_apply_kernel(const float* X, int64_t ldx, int64_t n, const float* t,
                                                        float* y, int nrhs)
{
    __shared__ float part[4][RES_ROWS * 8];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int64_t row0 = (int64_t)blockIdx.x * RES_ROWS;
    float acc[RES_ROWS][8];
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[i][q] = 0.0f;
    for (int64_t c = threadIdx.x; c < n; c += 256) {
        float tv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) tv[q] = q < nrhs ? t[c * nrhs + q] : 0.0f;
#pragma unroll
        for (int i = 0; i < RES_ROWS; ++i) {
            const float a = row0 + i < n ? X[(row0 + i) * ldx + c] : 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[i][q] = fmaf(a, tv[q], acc[i][q]);
        }
    }
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            float v = acc[i][q];
#pragma unroll
            for (int off = 32; off > 0; off /= 2) v += __shfl_xor(v, off);
            acc[i][q] = v;
        }
    if (lane < RES_ROWS * 8) {
        const int i = lane / 8, q = lane % 8;
        float v = 0.0f;
#pragma unroll
        for (int ii = 0; ii < RES_ROWS; ++ii)
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
                if (ii == i && qq == q) v = acc[ii][qq];
        part[wave][lane] = v;
    }
    __syncthreads();
    const int tt = threadIdx.x;
    if (tt < RES_ROWS * 8) {
        const int i = tt / 8, q = tt % 8;
        const float v = ((part[0][tt] + part[1][tt]) + part[2][tt]) + part[3][tt];
        if (q < nrhs && row0 + i < n) y[(row0 + i) * nrhs + q] = v;
    }
}

// X = P (the permutation of the factors: row i holds a one in column perm[i]).
__global__ __launch_bounds__(256) void perm_identity_kernel(float* X, int64_t ldx, int64_t n, const int32_t* perm)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t i = idx / n, j = idx % n;
    X[i * ldx + j] = j == perm[i] ? 1.0f : 0.0f;
}

// Per column q of an n x nrhs (ld) array: out[q] = sum_i a_iq^2 and
// out[8 + q] = max_i |a_iq| (one workgroup per column, fixed order).
__global__ __launch_bounds__(256) void colnorm_kernel(const double* a, int64_t ld, int64_t n, double* out, int64_t zb)
{
    a = zoff(a, zb);
    out = zoff(out, zb);

    __shared__ double ss[256], sm[256];
    const int q = blockIdx.x, tid = threadIdx.x;
    double s = 0.0, m = 0.0;
    for (int64_t i = tid; i < n; i += 256) {
        const double v = a[i * ld + q];
        s += v * v;
        m = fmax(m, fabs(v));
    }
    ss[tid] = s;
    sm[tid] = m;
    __syncthreads();
    for (int w = 128; w > 0; w /= 2) {
        if (tid < w) {
            ss[tid] += ss[tid + w];
            sm[tid] = fmax(sm[tid], sm[tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[q] = ss[0];
        out[8 + q] = sm[0];
    }
}

// x[i][q] (ld ldx) += scale[q] * d[i][q] (FP32 or FP64 correction, ld nrhs)
template <typename TD>
__global__ __launch_bounds__(256) void axpy_rows_kernel(double* x, int64_t ldx, const TD* d, int64_t n, int nrhs,
                                                        const double* scale, int64_t zb)
{
    x = zoff(x, zb);
    d = zoff(d, zb);
    scale = zoff(scale, zb);

    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    x[i * ldx + q] += scale[q] * (double)d[idx];
}

__global__ void scale_from_norms_kernel(const double* norms, double* scale, double* inv_scale, int nrhs, int64_t zb)
{
    norms = zoff(norms, zb);
    scale = zoff(scale, zb);
    inv_scale = zoff(inv_scale, zb);

    const int q = threadIdx.x;
    if (q >= nrhs) return;
    const double m = norms[8 + q];
    scale[q] = m > 0.0 ? m : 1.0;
    inv_scale[q] = m > 0.0 ? 1.0 / m : 1.0;
}

int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T>
struct FactorAux {
    // aux layout: perm (n int32, padded to 256 B), then per panel p
    // aux_blocks<T>() NB x NB blocks: invL, invU, and the chained solves'
    // links ML = inv(L_pp) L_p,p-1, MU = inv(U_pp) U_p,p+1, and for FP32
    // ML2 = inv(L_pp) L_p,p-2 and MU2 = inv(U_pp) U_p,p+2 (link_kernel)
    static constexpr int BLOCKS = aux_blocks<T>();
    static size_t perm_bytes(int64_t n) { return align_up((size_t)n * sizeof(int32_t), 256); }
    static size_t bytes(int64_t n)
    {
        return perm_bytes(n) + (size_t)div_up(n, NB) * BLOCKS * NB * NB * sizeof(T);
    }
    static int32_t* perm(void* aux) { return static_cast<int32_t*>(aux); }
    static T* invL(void* aux, int64_t n, int64_t p)
    {
        return reinterpret_cast<T*>(static_cast<char*>(aux) + perm_bytes(n)) + (size_t)p * BLOCKS * NB * NB;
    }
    static T* invU(void* aux, int64_t n, int64_t p) { return invL(aux, n, p) + NB * NB; }
    static T* linkL(void* aux, int64_t n, int64_t p) { return invL(aux, n, p) + 2 * NB * NB; }
    static T* linkU(void* aux, int64_t n, int64_t p) { return invL(aux, n, p) + 3 * NB * NB; }
};

// The links of the chained triangular solves: for block row p,
//   ML_p  = inv(L_pp) L_p,p-1  (p >= 1),   MU_p  = inv(U_pp) U_p,p+1  (p <= nblk - 2),
//   ML2_p = inv(L_pp) L_p,p-2  (p >= 2),   MU2_p = inv(U_pp) U_p,p+2  (p <= nblk - 3),
// NB x NB (ld NB), zero beyond the matrix and where the neighbour block does
// not exist.  With them a sweep's block step depends on its two nearest
// blocks only through x_p = c_p - M2_p x_(p-/+2) - M_p x_(p-/+1), c_p being
// formed before either arrives.  grid (nblk, 4): y = 0 forms ML, 1 MU, 2 ML2,
// 3 MU2; 256 lanes, each an 8 x 8 tile of the product, K staged through LDS
// in slices of 32 (FP64 factors: y = 0, 1 only).  blocks =
// FactorAux::invL(aux, n, 0) (the blocks of panel p start at
// p aux_blocks<T>() NB^2).
template <typename T>
__global__ __launch_bounds__(256) void link_kernel(const T* LU, int64_t lda, int64_t n, T* blocks, int64_t zb)
{
    LU = zoff(LU, zb);
    blocks = zoff(blocks, zb);

    constexpr int KC = 32;
    __shared__ T Is[NB * (KC + 1)];
    __shared__ T Bs[KC * NB];
    const int64_t p = blockIdx.x;
    const bool upper = blockIdx.y % 2 == 1;
    const int dist = blockIdx.y < 2 ? 1 : 2;
    const int64_t nblk = (n + NB - 1) / NB;
    T* blk = blocks + p * FactorAux<T>::BLOCKS * NB * NB;
    const T* inv = blk + (upper ? NB * NB : 0);
    T* out = blk + (2 + blockIdx.y) * NB * NB;
    const bool exists = upper ? p + dist < nblk : p >= dist;
    const int tid = threadIdx.x, ti = tid / 16, tj = tid % 16;
    T acc[8][8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[u][v] = T(0);
    if (exists) {
        const int64_t r0 = p * NB, c0 = (upper ? p + dist : p - dist) * NB;
        const int64_t nbr = min((int64_t)NB, n - r0), nbc = min((int64_t)NB, n - c0);
        for (int k0 = 0; k0 < NB; k0 += KC) {
            for (int idx = tid; idx < NB * KC; idx += 256) {
                const int i = idx / KC, k = idx % KC;
                Is[i * (KC + 1) + k] = inv[i * NB + k0 + k];
                const int kk = idx / NB, j = idx % NB;
                Bs[kk * NB + j] = (k0 + kk < nbr && j < nbc) ? LU[(r0 + k0 + kk) * lda + c0 + j] : T(0);
            }
            __syncthreads();
#pragma unroll 4
            for (int k = 0; k < KC; ++k) {
                T a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = Is[(ti * 8 + u) * (KC + 1) + k];
#pragma unroll
                for (int v = 0; v < 8; ++v) b[v] = Bs[k * NB + tj * 8 + v];
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int v = 0; v < 8; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) out[(ti * 8 + u) * NB + tj * 8 + v] = acc[u][v];
}

// ---------------------------------------------------------------- trailing update
// A22 -= L21 U12 is a plain GEMM on this library's MFMA kernels: FP32 and
// FP64 products on gemm_kernel (within a few percent of the vendor GEMM on
// these K = 256 shapes), and, with `split` workspace (FP32 factorizations
// with split-bf16 products), updates of 2048 or more rows and columns on the
// split-bf16 kernel (split3_update): FP32 operands split into bf16 hi and lo
// parts, three bf16 MFMA products per FP32 product, FP32 sums.  The FP32
// factors they give serve as FP64 refinement's preconditioner, whose
// convergence they leave unchanged on these matrices.

size_t split_ws_bytes(int64_t n) { return align_up(split3_workspace_bytes(n, n, NBO_SPLIT), 256); }
// the packed FP32 update's need for an update of M rows (M >= N, K <= NBO_SPLIT), within split_ws_bytes(n)
size_t split_ws_bytes_f(int64_t M) { return f32_workspace_bytes(M, M, NBO_SPLIT); }

// With `lower` (M == N; the symmetric factorization) only C's tiles on or
// below the diagonal are updated.
// With `packed` (FP32 products) the large updates run on the packed-operand
// MFMA kernel, which reaches a higher fraction of the FP32 MFMA peak than
// gemm_kernel on these shapes; `packed` and `split` share one workspace.
template <typename T>
int trailing_update(int64_t M, int64_t N, int64_t K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc,
                    void* split, size_t split_bytes, hipStream_t st, bool lower = false, void* packed = nullptr)
{
    if constexpr (std::is_same_v<T, float>) {
        if (split && M >= 2048 && N >= 256 && K <= NBO_SPLIT)
            return split3_update(M, N, K, A, lda, B, ldb, C, ldc, lower ? 1 : 0, split, split_bytes, st);
        if (packed && M >= 2048 && N >= 256 && K <= NBO_SPLIT)
            return f32_update(M, N, K, A, lda, B, ldb, C, ldc, lower ? 1 : 0, packed, split_ws_bytes_f(M), st);
    }
    if (lower) return gemm_lower_update<T>(M, K, A, lda, B, ldb, C, ldc, st);
    return gemm_launch<T>(M, N, K, T(-1), A, lda, B, ldb, T(1), C, ldc, nullptr, st);
}

// Symmetric factorization: U = D L^T.  Rows [r0, r0 + h) of U, columns
// [c0, c1): U[r][c] = U[r][r] L[c][r], the row block formed from the
// already-final columns of L below it (an LDS-tiled transpose: 32 x 32 tiles,
// coalesced reads of L's rows and writes of U's rows).
template <typename T>
__global__ __launch_bounds__(256) void sym_u_kernel(T* A, int64_t lda, int64_t r0, int h, int64_t c0, int64_t c1)
{
    __shared__ T t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int64_t cb = c0 + (int64_t)blockIdx.x * 32, rb = r0 + (int64_t)blockIdx.y * 32;
    // read L[cb + i][rb + tx] for i = ty, ty + 8, ...
    for (int i = ty; i < 32; i += 8) {
        const int64_t c = cb + i, r = rb + tx;
        t[i][tx] = (c < c1 && r < r0 + h) ? A[c * lda + r] : T(0);
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int64_t r = rb + i, c = cb + tx;
        if (r < r0 + h && c < c1) A[r * lda + c] = A[r * lda + r] * t[tx][i];
    }
}

template <typename T>
int sym_u_launch(T* A, int64_t lda, int64_t r0, int h, int64_t c0, int64_t c1, hipStream_t st)
{
    if (c1 <= c0 || h <= 0) return FL_OK;
    hipLaunchKernelGGL(sym_u_kernel<T>, dim3((unsigned)div_up(c1 - c0, 32), (unsigned)div_up(h, 32)), dim3(256), 0, st,
                       A, lda, r0, h, c0, c1);
    FL_LAUNCHED();
    return FL_OK;
}

// ---------------------------------------------------------------- look-ahead
// Orders of at least LOOKAHEAD_MIN factor with look-ahead (FL_LU_LOOKAHEAD=0
// turns it off, a measurement switch).
constexpr int64_t LOOKAHEAD_MIN = 4096;

bool sym_or_lu_lookahead(int64_t n)
{
    static const bool on = [] {
        const char* e = std::getenv("FL_LU_LOOKAHEAD");
        return !(e && *e == '0');
    }();
    return on && n >= LOOKAHEAD_MIN;
}

// A stream being captured into a graph keeps the single-stream order.
bool stream_capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// The look-ahead stream of (device, caller stream): created once, at the
// default priority, kept for the process.  Callers that share a stream share
// it too, which only serializes their panels (each call orders its work with
// its own events).
hipStream_t lookahead_stream(hipStream_t caller)
{
    static std::mutex mu;
    static std::vector<std::tuple<int, hipStream_t, hipStream_t>> pool;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    for (auto& [d, c, a] : pool)
        if (d == dev && c == caller) return a;
    hipStream_t aux = nullptr;
    if (hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) != hipSuccess) return nullptr;
    pool.emplace_back(dev, caller, aux);
    return aux;
}

// Events of one factorization (released when it ends; the device finishes
// with them on its own).
struct EventSet {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    int create()
    {
        for (auto& x : e) FL_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        return FL_OK;
    }
    ~EventSet()
    {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Workspace: the sub-panel backup (n x NB), flags, and for FP32 the two
// split-bf16 operand buffers of the trailing update.
template <typename T>
size_t getrf_ws_bytes(int64_t n)
{
    const size_t base = align_up((size_t)n * NB * sizeof(T), 256) + 256;
    return std::is_same_v<T, float> ? base + split_ws_bytes(n) : base;
}

// End of the asynchronous factorization: status = 0 when no multiplier
// exceeded 1 + 2^-20 (then the no-interchange pivots are partial pivoting's
// and the factors are those of the verified path), 1 otherwise; the row
// permutation is the identity.
__global__ void spec_finish_kernel(const unsigned int* lmax_bits, unsigned int* status, int32_t* perm, int64_t n, int64_t zb)
{
    lmax_bits = zoff(lmax_bits, zb);
    status = zoff(status, zb);
    perm = zoff(perm, zb);

    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = (int32_t)i;
    if (i == 0) {
        const float lmax = __uint_as_float(*lmax_bits);
        status[0] = lmax <= 1.0f + 1.0f / (1 << 20) ? 0u : 1u;
    }
}

// Blocked right-looking LU.  With status == nullptr (verified path) every
// sub-panel tries the three tiers in turn, reading the panel's largest
// multiplier back on the host; with a status word (asynchronous path) every
// sub-panel is factored without interchanges, the largest multiplier of the
// whole matrix is folded into one device word, and nothing is read back: the
// call only enqueues work, and *status says afterwards whether the result is
// valid (else A must be refactored from a copy on the verified path).
//
// With `sym` (asynchronous path only) A is taken as symmetric and factored
// as A = L D L^T without interchanges: only the lower triangle is read below
// the diagonal blocks, each row block of U is formed as D L^T from the
// finished columns of L (no U12 GEMMs), and trailing updates cover only the
// tiles on or below the diagonal, which halves the bulk of the work.  The
// factors are stored as fl_?getrf_async stores them (L below the diagonal,
// U = D L^T on and above it, the same aux), so the solves and the refinement
// take them unchanged; status is checked the same way.
template <typename T>
int getrf_impl(T* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info_host, int* fallback_panels,
               void* ws, size_t ws_bytes, hipStream_t st, unsigned int* status = nullptr, bool split_products = false,
               bool sym = false)
{
    const bool spec = status != nullptr;
    FL_CHECK(!sym || spec, "the symmetric factorization is asynchronous only");
    FL_CHECK(n >= 0, "negative n");
    FL_CHECK(lda >= n, "lda < n");
    if (n == 0) {
        if (spec) {
            hipLaunchKernelGGL(zero_u32_kernel, zgrid(dim3(1)), dim3(64), 0, st, status, (int64_t)1, batch().zb);
            FL_LAUNCHED();
        }
        return FL_OK;
    }
    FL_CHECK(A && ipiv && aux, "null pointer argument");
    if (ws_bytes < getrf_ws_bytes<T>(n) || !ws)
        return set_error(FL_EWORKSPACE, "getrf needs %zu workspace bytes", getrf_ws_bytes<T>(n));
    T* backup = static_cast<T*>(ws);
    char* tail = static_cast<char*>(ws) + align_up((size_t)n * NB * sizeof(T), 256);
    unsigned int* flag = reinterpret_cast<unsigned int*>(tail);      // largest |multiplier| of a panel
    int32_t* info_blk = reinterpret_cast<int32_t*>(tail + 4);        // zero pivot found by tier 1
    int32_t* info_d = reinterpret_cast<int32_t*>(tail + 64);
    // The split products run with NBO_SPLIT-column panels, and pay only where a
    // trailing update then reaches the library GEMM (n - NBO_SPLIT >= 2048).
    // Below that the factorization is the FP32-product one (256-column
    // panels, faster there).
    void* split = split_products && std::is_same_v<T, float> && n - NBO_SPLIT >= 2048 ? tail + 256 : nullptr;
    FL_CHECK(batch().count == 1 || (spec && !sym && !split),
             "a grouped factorization is the asynchronous one with FP32 products");
    // FP32 products: the large trailing updates on packed operands (FL_LU_F32_PACKED=0 turns it off, a
    // measurement switch)
    static const bool packed_on = [] {
        const char* e = std::getenv("FL_LU_F32_PACKED");
        return !(e && *e == '0');
    }();
    void* packed = std::is_same_v<T, float> && !split && packed_on && batch().count == 1 && n >= 2048 + 256
                       ? tail + 256 : nullptr;
    if (spec) {
        hipLaunchKernelGGL(zero_u32_kernel, zgrid(dim3(1)), dim3(64), 0, st, flag, (int64_t)1, batch().zb);
        FL_LAUNCHED();
    } else {
        FL_HIP(hipMemsetAsync(info_d, 0, sizeof(int32_t), st));
    }
    int fallbacks = 0;
    auto swap_cols = [&](int64_t k0, int nb, int64_t c0, int64_t c1) -> int {
        if (c1 <= c0 || spec) return FL_OK;   // the asynchronous path makes no interchanges
        hipLaunchKernelGGL(laswp_kernel<T>, dim3((unsigned)div_up(c1 - c0, 256)), dim3(256), 0, st, A, lda, k0, nb,
                           (const int32_t*)ipiv, c0, c1);
        FL_LAUNCHED();
        return FL_OK;
    };
    // Outer panels of nbo_w columns (the trailing update has K = nbo_w); each is
    // factored as sub-panels of NB columns.
    static const int64_t nbo_env = [] {
        const char* e = std::getenv("FL_LU_PANEL");   // measurement switch: outer panel width without the split
        const int v = e && *e ? std::atoi(e) : 0;
        return (int64_t)((v == 256 || v == 512) ? v : 0);
    }();
    // Large orders: 512-column outer panels (the trailing updates' K = 512
    // halves the passes over A22 and the FP32 updates' epilogues, which read C
    // back; FP64 7 % and FP32 products 2 % faster at n = 24,576, equal below
    // 12,000)
    const int64_t nbo_def = n >= 16384 ? 512 : NBO;
    const int64_t nbo_w = split ? NBO_SPLIT : (nbo_env ? nbo_env : nbo_def);
    // Look-ahead (asynchronous path, n >= LOOKAHEAD_MIN): while the caller's
    // stream updates the trailing matrix beyond the next outer panel, the next
    // panel (updated first) is factored on a second stream, so its diagonal
    // blocks and panel products overlap the bulk of the update.  Events order
    // the two streams; each touches disjoint parts of A in the meantime.
    const bool ahead = spec && sym_or_lu_lookahead(n) && batch().count == 1 && !stream_capturing(st);
    hipStream_t sp = st;
    EventSet ev;
    if (ahead) {
        if (int s = ev.create()) return s;
        sp = lookahead_stream(st);
        if (!sp) return set_error(FL_EHIP, "could not create the look-ahead stream");
        FL_HIP(hipEventRecord(ev.e[0], st));
        FL_HIP(hipStreamWaitEvent(sp, ev.e[0], 0));
    }
    auto factor_panel = [&](int64_t k) -> int {
        hipStream_t st = sp;   // the panel's work runs on the look-ahead stream when there is one
        const int64_t pe = std::min<int64_t>(n, k + nbo_w);
        for (int64_t ks = k; ks < pe; ks += NB) {
            const int nb = (int)std::min<int64_t>(NB, pe - ks);
            const int64_t m = n - ks;
            const int64_t p = ks / NB;
            T* Akk = A + ks * lda + ks;
            T* invL = FactorAux<T>::invL(aux, n, p);
            T* invU = FactorAux<T>::invU(aux, n, p);
            bool u12_done = false;   // the in-panel U12 already formed (grouped with L21)
            if (spec) {
                hipLaunchKernelGGL(diag_block_kernel<T>, zgrid(dim3(1)), dim3(256), 0, st, Akk, lda, nb, ks, ipiv, flag, invL,
                                   invU, batch().zb);
                FL_LAUNCHED();
                const int64_t rest0 = pe - ks - nb;
                if (sym && m > nb) {
                    // L21 = A21 inv(U11) (the lower part is up to date)
                    if (int s = gemm_launch<T>(m - nb, nb, nb, T(1), Akk + nb * lda, lda, invU, NB, T(0),
                                               Akk + nb * lda, lda, flag, st))
                        return s;
                    // the in-panel U12 = D11 L21^T
                    if (int s = sym_u_launch<T>(A, lda, ks, nb, ks + nb, pe, st)) return s;
                    u12_done = true;
                } else if (m > nb && rest0 > 0) {
                    // L21 = A21 inv(U11) and U12 = inv(L11) A12 (in-panel), disjoint, in one launch
                    if (int s = gemm_pair_launch<T>(m - nb, nb, nb, T(1), Akk + nb * lda, lda, invU, NB, T(0),
                                                    Akk + nb * lda, lda, flag, nb, rest0, nb, T(1), invL, NB, Akk + nb,
                                                    lda, T(0), Akk + nb, lda, st))
                        return s;
                    u12_done = true;
                } else if (m > nb) {
                    if (int s = gemm_launch<T>(m - nb, nb, nb, T(1), Akk + nb * lda, lda, invU, NB, T(0),
                                               Akk + nb * lda, lda, flag, st))
                        return s;
                }
            }
            // Tier 0: no interchanges; tier 1: interchanges inside the diagonal
            // block; tier 2: pivot search over every row below the diagonal.
            // A tier is accepted when no multiplier of the panel exceeds 1
            // (then its pivots are the ones partial pivoting chooses).
            if (!spec)
                FL_HIP(hipMemcpy2DAsync(backup, nb * sizeof(T), Akk, lda * sizeof(T), nb * sizeof(T), m,
                                        hipMemcpyDeviceToDevice, st));
            for (int tier = 0; !spec; ++tier) {
                if (tier == 2) {
                    ++fallbacks;
                    for (int64_t c = ks; c < ks + nb; ++c) {
                        hipLaunchKernelGGL(pivot_col_kernel<T>, dim3(1), dim3(1024), 0, st, A, lda, n, ks, nb, c, ipiv,
                                           info_d);
                        FL_LAUNCHED();
                        if (c + 1 < ks + nb && c + 1 < n) {
                            hipLaunchKernelGGL(panel_update_kernel<T>, dim3((unsigned)div_up(n - c - 1, 64)), dim3(256),
                                               0, st, A, lda, n, ks, nb, c);
                            FL_LAUNCHED();
                        }
                    }
                    hipLaunchKernelGGL(trtri_kernel<T>, dim3(2, NB / TR_COLS), dim3(128), 0, st, (const T*)Akk, lda,
                                       nb, invL, invU);
                    FL_LAUNCHED();
                    break;
                }
                FL_HIP(hipMemsetAsync(flag, 0, 2 * sizeof(unsigned int), st));  // flag and info_blk
                if (tier == 0)
                    hipLaunchKernelGGL(diag_block_kernel<T>, zgrid(dim3(1)), dim3(256), 0, st, Akk, lda, nb, ks, ipiv, flag,
                                       invL, invU, batch().zb);
                else
                    hipLaunchKernelGGL(diag_lu_kernel<T>, dim3(1), dim3(256), 0, st, Akk, lda, nb, ks, ipiv, info_blk,
                                       invL, invU);
                FL_LAUNCHED();
                // L21 = A21 inv(U11), in place, with max |L21| folded into the flag
                if (m > nb)
                    if (int s = gemm_launch<T>(m - nb, nb, nb, T(1), Akk + nb * lda, lda, invU, NB, T(0),
                                               Akk + nb * lda, lda, flag, st))
                        return s;
                unsigned int bits = 0;
                FL_HIP(hipMemcpyAsync(&bits, flag, sizeof(bits), hipMemcpyDeviceToHost, st));
                FL_HIP(hipStreamSynchronize(st));
                float lmax;
                std::memcpy(&lmax, &bits, sizeof(float));
                if (lmax <= 1.0f + 1.0f / (1 << 20)) {
                    if (tier == 1) {
                        hipLaunchKernelGGL(fold_info_kernel, dim3(1), dim3(1), 0, st, (const int32_t*)info_blk, info_d);
                        FL_LAUNCHED();
                    }
                    break;
                }
                // restore the sub-panel (and a clean info) and try the next tier
                FL_HIP(hipMemcpy2DAsync(Akk, lda * sizeof(T), backup, nb * sizeof(T), nb * sizeof(T), m,
                                        hipMemcpyDeviceToDevice, st));
            }
            // this sub-panel's interchanges on the rest of the outer panel
            if (int s = swap_cols(ks, nb, k, ks)) return s;
            if (int s = swap_cols(ks, nb, ks + nb, pe)) return s;
            // update the outer panel's remaining columns
            const int64_t rest = pe - ks - nb;
            if (rest > 0) {
                T* A12 = Akk + nb;
                if (!u12_done)
                    if (int s = gemm_launch<T>(nb, rest, nb, T(1), invL, NB, A12, lda, T(0), A12, lda, nullptr, st))
                        return s;
                if (m > nb)
                    if (int s = gemm_launch<T>(m - nb, rest, nb, T(-1), Akk + nb * lda, lda, A12, lda, T(1),
                                               Akk + nb * lda + nb, lda, nullptr, st))
                        return s;
            }
        }
        return FL_OK;
    };
    if (int s = factor_panel(0)) return s;
    for (int64_t k = 0; k < n; k += nbo_w) {
        const int64_t pe = std::min<int64_t>(n, k + nbo_w);
        if (ahead) {   // the panel's factors, before anything reads them
            FL_HIP(hipEventRecord(ev.e[1], sp));
            FL_HIP(hipStreamWaitEvent(st, ev.e[1], 0));
        }
        const int nbo = (int)(pe - k);
        // the outer panel's interchanges to its left and right
        if (int s = swap_cols(k, nbo, 0, k)) return s;
        if (int s = swap_cols(k, nbo, pe, n)) return s;
        const int64_t nr = n - pe;
        if (nr <= 0) break;
        const size_t sb = split ? split_ws_bytes(n) : 0;
        if (sym) {
            // U12 = D L21^T for the whole outer panel
            if (int s = sym_u_launch<T>(A, lda, k, nbo, pe, n, st)) return s;
        } else {
            // U12 = inv(L11) A12 block row by block row
            for (int64_t ks = k; ks < pe; ks += NB) {
                const int nb = (int)std::min<int64_t>(NB, pe - ks);
                T* Ar = A + ks * lda + pe;
                if (int s = gemm_launch<T>(nb, nr, nb, T(1), FactorAux<T>::invL(aux, n, ks / NB), NB, Ar, lda, T(0),
                                           Ar, lda, nullptr, st))
                    return s;
                const int64_t below = pe - ks - nb;
                if (below > 0)
                    if (int s = gemm_launch<T>(below, nr, nb, T(-1), A + (ks + nb) * lda + ks, lda, Ar, lda, T(1),
                                               A + (ks + nb) * lda + pe, lda, nullptr, st))
                        return s;
            }
        }
        T* L21 = A + pe * lda + k;   // rows pe .. n of the panel's L
        const T* U12 = A + k * lda;  // row block k .. pe of U (column offset added below)
        if (!ahead) {
            // trailing update with K = nbo (symmetric: the tiles on or below the diagonal)
            if (int s = trailing_update<T>(nr, nr, nbo, L21, lda, U12 + pe, lda, A + pe * lda + pe, lda, split, sb, st, sym,
                                           packed))
                return s;
            if (int s = factor_panel(pe)) return s;
            continue;
        }
        // (a) the next panel's columns pe .. pe2, all rows from pe; then that
        // panel is factored on the look-ahead stream while (b) updates the rest
        const int64_t pe2 = std::min<int64_t>(n, pe + nbo_w);
        if (int s = trailing_update<T>(nr, pe2 - pe, nbo, L21, lda, U12 + pe, lda, A + pe * lda + pe, lda, split, sb, st,
                                       false, packed))
            return s;
        FL_HIP(hipEventRecord(ev.e[2], st));
        FL_HIP(hipStreamWaitEvent(sp, ev.e[2], 0));
        if (int s = factor_panel(pe)) return s;
        if (pe2 < n) {
            if (sym) {
                const int64_t r2 = n - pe2;
                if (int s = trailing_update<T>(r2, r2, nbo, A + pe2 * lda + k, lda, U12 + pe2, lda, A + pe2 * lda + pe2,
                                               lda, split, sb, st, true, packed))
                    return s;
            } else if (int s = trailing_update<T>(nr, n - pe2, nbo, L21, lda, U12 + pe2, lda, A + pe * lda + pe2, lda,
                                                  split, sb, st, false, packed)) {
                return s;
            }
        }
    }
    // the chained solves' links, from the finished factors
    hipLaunchKernelGGL(link_kernel<T>, zgrid(dim3((unsigned)div_up(n, NB), aux_blocks<T>() - 2)), dim3(256), 0, st,
                       (const T*)A, lda, n,
                       FactorAux<T>::invL(aux, n, 0), batch().zb);
    FL_LAUNCHED();
    if (spec) {
        hipLaunchKernelGGL(spec_finish_kernel, zgrid(dim3((unsigned)div_up(n, 256))), dim3(256), 0, st, (const unsigned*)flag,
                           status, FactorAux<T>::perm(aux), n, batch().zb);
        FL_LAUNCHED();
        return FL_OK;
    }
    // fold the interchanges into a row permutation (host; ipiv is n int32)
    std::vector<int32_t> piv(n), perm(n);
    FL_HIP(hipMemcpyAsync(piv.data(), ipiv, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    int32_t info = 0;
    FL_HIP(hipMemcpyAsync(&info, info_d, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[piv[i]]);
    FL_HIP(hipMemcpyAsync(FactorAux<T>::perm(aux), perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FL_HIP(hipStreamSynchronize(st));
    if (info_host) *info_host = info;
    if (fallback_panels) *fallback_panels = fallbacks;
    return FL_OK;
}

template <typename T>
size_t getrs_ws_bytes(int64_t n, int64_t nrhs)
{
    return align_up((size_t)n * nrhs * sizeof(T), 256);
}

// Words of the chained solves' sync area (two sweeps).
size_t chain_sync_bytes(int64_t n) { return align_up((size_t)2 * (div_up(n, NB) + 2) * sizeof(unsigned), 256); }

// Two lanes per row (256-lane workgroups).  Four lanes per row (512-lane
// workgroups, half the products per lane) measured slower: 11.0 against
// 8.4 us per FP32 block step at n = 5,400, the wider barriers and staging
// costing more than the shorter products save.
constexpr int CH_TPR = 2;

template <typename T, bool LOWER>
void launch_chain(const T* LU, int64_t lda, int64_t n, const T* invs, const T* b, T* x, int nrhs, unsigned* sync,
                  hipStream_t st)
{
    const dim3 grid((unsigned)div_up(n, NB)), blk(NB * CH_TPR);
    if (nrhs <= 1)
        hipLaunchKernelGGL((trsm_chain_kernel<T, LOWER, 1, CH_TPR>), zgrid(grid), blk, 0, st, LU, lda, n, invs, b, x, nrhs, sync, batch().zb);
    else if (nrhs <= 2)
        hipLaunchKernelGGL((trsm_chain_kernel<T, LOWER, 2, CH_TPR>), zgrid(grid), blk, 0, st, LU, lda, n, invs, b, x, nrhs, sync, batch().zb);
    else if (nrhs <= 4)
        hipLaunchKernelGGL((trsm_chain_kernel<T, LOWER, 4, CH_TPR>), zgrid(grid), blk, 0, st, LU, lda, n, invs, b, x, nrhs, sync, batch().zb);
    else if (nrhs <= 6)   // the six rigid-body modes
        hipLaunchKernelGGL((trsm_chain_kernel<T, LOWER, 6, CH_TPR>), zgrid(grid), blk, 0, st, LU, lda, n, invs, b, x, nrhs, sync, batch().zb);
    else hipLaunchKernelGGL((trsm_chain_kernel<T, LOWER, 8, CH_TPR>), zgrid(grid), blk, 0, st, LU, lda, n, invs, b, x, nrhs, sync, batch().zb);
}

// Solves in place on a contiguous n x nrhs buffer `y` that already holds the
// permuted right-hand side; tmp is another n x nrhs buffer, sync holds
// chain_sync_bytes(n).  Result in y.
template <typename T>
int solve_permuted(const T* LU, int64_t n, int64_t lda, void* aux, T* y, T* tmp, int nrhs, unsigned* sync,
                   hipStream_t st)
{
    const int64_t words = (int64_t)(chain_sync_bytes(n) / sizeof(unsigned));
    hipLaunchKernelGGL(zero_u32_kernel, zgrid(dim3((unsigned)div_up(words, 256))), dim3(256), 0, st, sync, words, batch().zb);
    FL_LAUNCHED();
    const T* invs = FactorAux<T>::invL(aux, n, 0);
    const int64_t half = div_up(n, NB) + 2;
    const int64_t len = n * nrhs;
    const dim3 fg((unsigned)div_up(len, 256));
    hipLaunchKernelGGL(fill_sentinel_kernel<T>, zgrid(fg), dim3(256), 0, st, tmp, len, batch().zb);
    FL_LAUNCHED();
    launch_chain<T, true>(LU, lda, n, invs, (const T*)y, tmp, nrhs, sync, st);
    FL_LAUNCHED();
    hipLaunchKernelGGL(fill_sentinel_kernel<T>, zgrid(fg), dim3(256), 0, st, y, len, batch().zb);   // y was the forward sweep's input
    FL_LAUNCHED();
    launch_chain<T, false>(LU, lda, n, invs, (const T*)tmp, y, nrhs, sync + half, st);
    FL_LAUNCHED();
    return FL_OK;
}

template <typename T>
int getrs_impl(const T* LU, int64_t n, int64_t lda, const void* aux_c, T* B, int64_t nrhs, int64_t ldb, void* ws,
               size_t ws_bytes, hipStream_t st)
{
    FL_CHECK(n >= 0 && nrhs >= 0, "negative size");
    FL_CHECK(nrhs <= 8, "at most 8 right-hand sides per call");
    FL_CHECK(ldb >= nrhs, "ldb < nrhs");
    if (n == 0 || nrhs == 0) return FL_OK;
    const size_t one = getrs_ws_bytes<T>(n, nrhs);
    const size_t need = 2 * one + chain_sync_bytes(n);
    if (ws_bytes < need || !ws) return set_error(FL_EWORKSPACE, "getrs needs %zu workspace bytes", need);
    void* aux = const_cast<void*>(aux_c);
    T* y = static_cast<T*>(ws);
    T* tmp = reinterpret_cast<T*>(static_cast<char*>(ws) + one);
    unsigned* sync = reinterpret_cast<unsigned*>(static_cast<char*>(ws) + 2 * one);
    const unsigned g = (unsigned)div_up(n * nrhs, 256);
    hipLaunchKernelGGL((gather_rows_kernel<T, T>), zgrid(dim3(g)), dim3(256), 0, st, (const T*)B, ldb, y, n, (int)nrhs,
                       (const int32_t*)FactorAux<T>::perm(aux), (const double*)nullptr, batch().zb);
    FL_LAUNCHED();
    if (int s = solve_permuted<T>(LU, n, lda, aux, y, tmp, (int)nrhs, sync, st)) return s;
    hipLaunchKernelGGL((copy_rows_kernel<T, T>), zgrid(dim3(g)), dim3(256), 0, st, (const T*)y, B, ldb, n, (int)nrhs, batch().zb);
    FL_LAUNCHED();
    return FL_OK;
}

// Zeroes n doubles (the fixed-step refinement clears x with a kernel, not a
// memset, so that it can be captured in a graph).
__global__ void zero_f64_kernel(double* p, int64_t n, int64_t zb)
{
    p = zoff(p, zb);

    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Final relative residuals of the fixed-step refinement, on the device:
// ||r||_2 / ||b||_2 per column (||r||_2 when b = 0), from the squared norms;
// NaN when a chained solve's block wait timed out (e0, e1: its flags).
__global__ void finish_resid_kernel(const double* norms, const double* bnorms, const unsigned* e0, const unsigned* e1,
                                    double* out, int q, int64_t zb)
{
    norms = zoff(norms, zb);
    bnorms = zoff(bnorms, zb);
    e0 = zoff(e0, zb);
    e1 = zoff(e1, zb);
    out = zoff(out, zb);

    const int c = threadIdx.x;
    if (c >= q) return;
    const double bn = sqrt(bnorms[c]);
    double rel = bn > 0 ? sqrt(norms[c]) / bn : sqrt(norms[c]);
    if ((e0 && *e0) || (e1 && *e1)) rel = __builtin_nan("");
    out[c] = rel;
}

size_t refine_ws_bytes(int64_t n, int64_t nrhs)
{
    return 3 * align_up((size_t)n * nrhs * sizeof(double), 256) + 2 * align_up((size_t)n * nrhs * sizeof(float), 256) +
           1024 + chain_sync_bytes(n);
}

}  // namespace

int refine_impl(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux_c, int64_t n,
                const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
                double* resid_out, const double* frame, void* ws, size_t ws_bytes, hipStream_t st,
                const float* inv32 = nullptr, int64_t ldi = 0, double* resid_dev = nullptr)
{
    // With resid_dev the refinement runs exactly max_iter corrections with no
    // host synchronization and no memset (capturable in a HIP graph) and
    // writes the final relative residuals to resid_dev on the device.
    const bool fixed = resid_dev != nullptr;
    FL_CHECK(batch().count == 1 || (fixed && !frame && !inv32), "grouped refinement takes fixed steps on LU factors");
    FL_CHECK(!inv32 || ldi >= n, "ldi < n");
    FL_CHECK(!fixed || max_iter >= 0, "negative step count");
    FL_CHECK(n >= 0 && nrhs >= 1 && nrhs <= 8, "need 1 <= nrhs <= 8");
    FL_CHECK(ldb >= nrhs, "ldb < nrhs");
    FL_CHECK(!frame || n % 3 == 0, "a frame needs n = 3 x points");
    Frame3 R{};
    if (frame) {
        for (int k = 0; k < 9; ++k) {
            R.r[k] = frame[k];
            FL_CHECK(std::isfinite(R.r[k]), "frame entries must be finite");
        }
    }
    if (n == 0) return FL_OK;
    if (ws_bytes < refine_ws_bytes(n, nrhs) || !ws)
        return set_error(FL_EWORKSPACE, "refine needs %zu workspace bytes", refine_ws_bytes(n, nrhs));
    void* aux = const_cast<void*>(aux_c);
    char* w = static_cast<char*>(ws);
    const size_t d8 = align_up((size_t)n * nrhs * sizeof(double), 256);
    const size_t d4 = align_up((size_t)n * nrhs * sizeof(float), 256);
    double* r = reinterpret_cast<double*>(w);
    double* xc = reinterpret_cast<double*>(w + d8);  // contiguous x (ld nrhs)
    float* y = reinterpret_cast<float*>(w + 3 * d8);
    float* tmp = reinterpret_cast<float*>(w + 3 * d8 + d4);
    double* small = reinterpret_cast<double*>(w + 3 * d8 + 2 * d4);  // norms[16], scale[8], inv_scale[8], bnorm[16]
    double* norms = small;
    double* scale = small + 16;
    double* inv_scale = small + 24;
    double* bnorms = small + 32;
    unsigned* sync = reinterpret_cast<unsigned*>(w + 3 * d8 + 2 * d4 + 1024);
    const unsigned g = (unsigned)div_up(n * nrhs, 256);
    const int q = (int)nrhs;
    const bool pair = reinterpret_cast<uintptr_t>(A64) % 16 == 0 && lda64 % 2 == 0;

    // r = b (contiguous, ld nrhs) and ||b|| per column
    hipLaunchKernelGGL((gather_rows_kernel<double, double>), zgrid(dim3(g)), dim3(256), 0, st, b, ldb, r, n, q,
                       (const int32_t*)nullptr, (const double*)nullptr, batch().zb);
    FL_LAUNCHED();
    hipLaunchKernelGGL(colnorm_kernel, zgrid(dim3(q)), dim3(256), 0, st, (const double*)r, (int64_t)nrhs, n, bnorms, batch().zb);
    FL_LAUNCHED();
    if (fixed) {
        hipLaunchKernelGGL(zero_f64_kernel, zgrid(dim3(g)), dim3(256), 0, st, xc, n * nrhs, batch().zb);
        FL_LAUNCHED();
    } else {
        FL_HIP(hipMemsetAsync(xc, 0, (size_t)n * nrhs * sizeof(double), st));
    }
    // small = norms[16], scale[8], inv_scale[8], bnorms[16]: one read-back
    // brings the residual's and the right-hand side's norms to the host
    std::vector<double> hs(48);
    const double* hr = hs.data();
    const double* hb = hs.data() + 32;
    int it = 0;
    double worst = 0.0;
    for (;;) {
        // r holds the current residual b - A x.  Before the first correction
        // x = 0 and r = b, whose norms are already known (relative residual 1),
        // so the host looks only from the first correction on.
        if (it > 0 || max_iter <= 0) {
            hipLaunchKernelGGL(colnorm_kernel, zgrid(dim3(q)), dim3(256), 0, st, (const double*)r, (int64_t)nrhs, n, norms, batch().zb);
            FL_LAUNCHED();
            if (fixed) {
                if (it >= max_iter) break;   // the norms stay on the device
            } else {
                FL_HIP(hipMemcpyAsync(hs.data(), small, 48 * sizeof(double), hipMemcpyDeviceToHost, st));
                FL_HIP(hipStreamSynchronize(st));
                worst = 0.0;
                for (int c = 0; c < q; ++c) {
                    const double bn = std::sqrt(hb[c]);
                    const double rel = bn > 0 ? std::sqrt(hr[c]) / bn : std::sqrt(hr[c]);
                    if (resid_out) resid_out[c] = rel;
                    worst = std::max(worst, rel);
                }
                if ((it > 0 && worst <= tol) || it >= max_iter) break;
            }
        }
        // d = LU32^{-1} (r / s) s, with s = max |r| per column so FP32 sees O(1) values
        hipLaunchKernelGGL(scale_from_norms_kernel, zgrid(dim3(1)), dim3(8), 0, st, (const double*)(it == 0 ? bnorms : norms),
                           scale, inv_scale, q, batch().zb);
        // with an explicit inverse (which includes the row permutation) the
        // scaled residual goes to tmp unpermuted and y = inv32 tmp
        const int32_t* perm = inv32 ? nullptr : (const int32_t*)FactorAux<float>::perm(aux);
        float* yin = inv32 ? tmp : y;
        if (frame)
            hipLaunchKernelGGL(gather_rot_kernel, dim3(g), dim3(256), 0, st, (const double*)r, yin, n, q, perm,
                               (const double*)inv_scale, R);
        else
            hipLaunchKernelGGL((gather_rows_kernel<double, float>), zgrid(dim3(g)), dim3(256), 0, st, (const double*)r,
                               (int64_t)nrhs, yin, n, q, perm, (const double*)inv_scale, batch().zb);
        FL_LAUNCHED();
        if (inv32) {
            hipLaunchKernelGGL(inv_apply_kernel, dim3((unsigned)div_up(n, RES_ROWS)), dim3(256), 0, st, inv32, ldi, n,
                               (const float*)tmp, y, q);
            FL_LAUNCHED();
        } else if (int s = solve_permuted<float>(LU32, n, lda32, aux, y, tmp, q, sync, st)) {
            return s;
        }
        if (frame)
            hipLaunchKernelGGL(axpy_rot_kernel, dim3(g), dim3(256), 0, st, xc, (const float*)y, n, q,
                               (const double*)scale, R);
        else
            hipLaunchKernelGGL(axpy_rows_kernel<float>, zgrid(dim3(g)), dim3(256), 0, st, xc, (int64_t)nrhs, (const float*)y,
                               n, q, (const double*)scale, batch().zb);
        FL_LAUNCHED();
        ++it;
        if (pair)
            residual_launch<true>(q, zgrid(dim3((unsigned)div_up(n, RES_ROWS))), st, A64, lda64, n, b, ldb,
                                  (const double*)xc, r, batch().zb);
        else
            residual_launch<false>(q, zgrid(dim3((unsigned)div_up(n, RES_ROWS))), st, A64, lda64, n, b, ldb,
                                   (const double*)xc, r, batch().zb);
        FL_LAUNCHED();
    }
    hipLaunchKernelGGL((copy_rows_kernel<double, double>), zgrid(dim3(g)), dim3(256), 0, st, (const double*)xc, x, ldb, n, q, batch().zb);
    FL_LAUNCHED();
    if (fixed) {
        const int64_t half = div_up(n, NB) + 2;
        const bool waits = !inv32 && it > 0;
        hipLaunchKernelGGL(finish_resid_kernel, zgrid(dim3(1)), dim3(64), 0, st, (const double*)norms, (const double*)bnorms,
                           (const unsigned*)(waits ? sync + 1 : nullptr),
                           (const unsigned*)(waits ? sync + half + 1 : nullptr), resid_dev, q, batch().zb);
        FL_LAUNCHED();
        if (iters_out) *iters_out = it;
        return FL_OK;
    }
    unsigned sync_err[2] = {0u, 0u};
    if (!inv32 && it > 0) {   // the chained solves' wait flags (an explicit inverse has none)
        const int64_t half = div_up(n, NB) + 2;
        FL_HIP(hipMemcpyAsync(&sync_err[0], sync + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        FL_HIP(hipMemcpyAsync(&sync_err[1], sync + half + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    }
    FL_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = it;
    if (sync_err[0] | sync_err[1]) return set_error(FL_EHIP, "triangular solve: a block wait timed out");
    if (!(worst <= tol)) return set_error(FL_ENOCONV, "refinement reached %.3e after %d steps (tol %.1e)", worst, it, tol);
    return FL_OK;
}

// Explicit inverse from the FP32 factors of P A = L U: X = U^{-1} L^{-1} P,
// by blocked substitution on X = P.  Forward, block row p = 0, 1, ...:
//   X_p = inv(L_pp) X_p,  X_below -= L_below,p X_p;
// backward, p = last, ..., 0:
//   X_p = inv(U_pp) X_p,  X_above -= U_above,p X_p,
// with the inverted diagonal blocks kept in aux and the LU's update GEMMs
// (2 n^3 flop).  The refinement then corrects with one pass over X
// (inv_apply_kernel) instead of two chained triangular sweeps, which pays
// when one factorization serves many solves.
int getri_impl(const float* LU, int64_t n, int64_t lda, const void* aux_c, float* X, int64_t ldx, hipStream_t st)
{
    FL_CHECK(n >= 0, "negative n");
    FL_CHECK(lda >= n && ldx >= n, "leading dimension < n");
    if (n == 0) return FL_OK;
    FL_CHECK(LU && aux_c && X, "null pointer argument");
    void* aux = const_cast<void*>(aux_c);
    hipLaunchKernelGGL(perm_identity_kernel, dim3((unsigned)div_up(n * n, 256)), dim3(256), 0, st, X, ldx, n,
                       (const int32_t*)FactorAux<float>::perm(aux));
    FL_LAUNCHED();
    const int64_t nblk = div_up(n, NB);
    for (int64_t p = 0; p < nblk; ++p) {
        const int64_t k0 = p * NB;
        const int nb = (int)std::min<int64_t>(NB, n - k0);
        float* Xp = X + k0 * ldx;
        // in place: M = nb <= 128 rows, C on B's memory
        if (int s = gemm_launch<float>(nb, n, nb, 1.0f, FactorAux<float>::invL(aux, n, p), NB, Xp, ldx, 0.0f, Xp, ldx,
                                       nullptr, st))
            return s;
        const int64_t below = n - k0 - nb;
        if (below > 0)
            if (int s = trailing_update<float>(below, n, nb, LU + (k0 + nb) * lda + k0, lda, Xp, ldx, Xp + nb * ldx, ldx,
                                               nullptr, 0, st))
                return s;
    }
    for (int64_t p = nblk - 1; p >= 0; --p) {
        const int64_t k0 = p * NB;
        const int nb = (int)std::min<int64_t>(NB, n - k0);
        float* Xp = X + k0 * ldx;
        if (int s = gemm_launch<float>(nb, n, nb, 1.0f, FactorAux<float>::invU(aux, n, p), NB, Xp, ldx, 0.0f, Xp, ldx,
                                       nullptr, st))
            return s;
        if (k0 > 0)
            if (int s = trailing_update<float>(k0, n, nb, LU + k0, lda, Xp, ldx, X, ldx, nullptr, 0, st)) return s;
    }
    return FL_OK;
}

namespace {

// The identity padding of an order-n solve held in buffers of order n_pad
// (row-major, leading dimension n_pad): rows and columns n..n_pad-1 of K64
// and K32 (either may be null) become the identity's, and b (n_pad x nrhs)
// gets rhs (n x nrhs) over zeros.  n is read on the device (a double at
// n_dev), so that one captured graph serves every n <= n_pad.  One
// workgroup per row; the leading n x n block is left alone.
__global__ __launch_bounds__(256) void pad_solve_kernel(double* K64, float* K32, int64_t n_pad, const double* n_dev,
                                                        const double* rhs, double* b, int nrhs, int64_t zb)
{
    K64 = zoff(K64, zb);
    K32 = zoff(K32, zb);
    n_dev = zoff(n_dev, zb);
    rhs = zoff(rhs, zb);
    b = zoff(b, zb);

    const int64_t n = std::min<int64_t>(std::max<int64_t>((int64_t)n_dev[0], 0), n_pad);
    const int64_t i = blockIdx.x;
    double* r64 = K64 ? K64 + i * n_pad : nullptr;
    float* r32 = K32 ? K32 + i * n_pad : nullptr;
    for (int64_t c = (i < n ? n : 0) + threadIdx.x; c < n_pad; c += blockDim.x) {
        const double v = c == i ? 1.0 : 0.0;
        if (r64) r64[c] = v;
        if (r32) r32[c] = (float)v;
    }
    if ((int)threadIdx.x < nrhs) b[i * nrhs + threadIdx.x] = i < n ? rhs[i * nrhs + threadIdx.x] : 0.0;
}

}  // namespace

int pad_solve_impl(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                   int64_t nrhs, hipStream_t st)
{
    FL_CHECK(n_pad >= 0, "negative order");
    FL_CHECK(nrhs >= 1 && nrhs <= 8, "need 1 <= nrhs <= 8");
    if (n_pad == 0) return FL_OK;
    FL_CHECK(n_dev && rhs && b, "null pointer argument");
    hipLaunchKernelGGL(pad_solve_kernel, zgrid(dim3((unsigned)n_pad)), dim3(256), 0, st, K64, K32, n_pad, n_dev, rhs, b,
                       (int)nrhs, batch().zb);
    FL_LAUNCHED();
    return FL_OK;
}

size_t getrf_aux_bytes(int64_t n, int precision)
{
    return precision == FL_FP64 ? FactorAux<double>::bytes(n) : FactorAux<float>::bytes(n);
}
size_t getrf_workspace_bytes(int64_t n, int precision)
{
    return precision == FL_FP64 ? getrf_ws_bytes<double>(n) : getrf_ws_bytes<float>(n);
}
size_t getrs_workspace_bytes(int64_t n, int64_t nrhs, int precision)
{
    return 2 * (precision == FL_FP64 ? getrs_ws_bytes<double>(n, nrhs) : getrs_ws_bytes<float>(n, nrhs)) +
           chain_sync_bytes(n);
}
size_t refine_workspace_bytes(int64_t n, int64_t nrhs) { return refine_ws_bytes(n, nrhs); }

int sgetrf(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, int products, void* ws,
           size_t wsb, hipStream_t st)
{
    FL_CHECK(products == FL_LU_FP32 || products == FL_LU_BF16X3, "unknown products %d", products);
    return getrf_impl<float>(A, n, lda, ipiv, aux, info, fb, ws, wsb, st, nullptr, products == FL_LU_BF16X3);
}
int dgetrf(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, void* ws, size_t wsb,
           hipStream_t st)
{
    return getrf_impl<double>(A, n, lda, ipiv, aux, info, fb, ws, wsb, st);
}
int sgetrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                 void* ws, size_t wsb, hipStream_t st)
{
    FL_CHECK(status, "null status");
    FL_CHECK(products == FL_LU_FP32 || products == FL_LU_BF16X3, "unknown products %d", products);
    return getrf_impl<float>(A, n, lda, ipiv, aux, nullptr, nullptr, ws, wsb, st, status, products == FL_LU_BF16X3);
}
int dgetrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, void* ws,
                 size_t wsb, hipStream_t st)
{
    FL_CHECK(status, "null status");
    return getrf_impl<double>(A, n, lda, ipiv, aux, nullptr, nullptr, ws, wsb, st, status);
}
int ssytrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                 void* ws, size_t wsb, hipStream_t st)
{
    FL_CHECK(status, "null status");
    FL_CHECK(products == FL_LU_FP32 || products == FL_LU_BF16X3, "unknown products %d", products);
    return getrf_impl<float>(A, n, lda, ipiv, aux, nullptr, nullptr, ws, wsb, st, status, products == FL_LU_BF16X3,
                             true);
}
int dsytrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, void* ws,
                 size_t wsb, hipStream_t st)
{
    FL_CHECK(status, "null status");
    return getrf_impl<double>(A, n, lda, ipiv, aux, nullptr, nullptr, ws, wsb, st, status, false, true);
}
int sgetrs(const float* LU, int64_t n, int64_t lda, const void* aux, float* B, int64_t nrhs, int64_t ldb, void* ws,
           size_t wsb, hipStream_t st)
{
    return getrs_impl<float>(LU, n, lda, aux, B, nrhs, ldb, ws, wsb, st);
}
int dgetrs(const double* LU, int64_t n, int64_t lda, const void* aux, double* B, int64_t nrhs, int64_t ldb, void* ws,
           size_t wsb, hipStream_t st)
{
    return getrs_impl<double>(LU, n, lda, aux, B, nrhs, ldb, ws, wsb, st);
}

}  // namespace fl
