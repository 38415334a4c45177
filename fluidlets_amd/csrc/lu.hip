// Blocked right-looking LU with partial pivoting, triangular solves and FP64
// iterative refinement, for the dense resistance matrices.
//
// Panel step k (width nb <= NB = 128), in row-major storage:
//   1. the nb x nb diagonal block is factored by GEPP in LDS by one workgroup,
//      its pivot search restricted to the block's rows;
//   2. inv(L11) and inv(U11) are formed (one workgroup each) and kept in the
//      factor's aux buffer for the triangular solves;
//   3. L21 = A21 inv(U11) on MFMA, in place; the GEMM epilogue records
//      max |L21|.  If it is <= 1 every pivot chosen in step 1 is also the
//      pivot partial pivoting over the whole column would choose (all
//      multipliers below the block are bounded by the block's pivot), so the
//      factorization is exactly GEPP.  Otherwise the panel is restored from a
//      copy and refactored column by column with the pivot search over all
//      rows below the diagonal (the general path);
//   4. the panel's row interchanges are applied left and right of it;
//   5. U12 = inv(L11) A12 on MFMA, in place;
//   6. A22 -= L21 U12 on MFMA (FP32 or FP64 accumulation).
// The LAPACK pivot vector (ipiv[r] = row swapped with r) is also folded into a
// row permutation stored with the factor, so solves gather the right-hand
// side in one pass.
//
// Replaces LAPACK getrf/getrs behind scipy.linalg.solve
// (src/fluidlets/core/solve.py:66) and numpy.linalg.solve
// (src/fluidlets/swimming.py:101).
#include "fl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace fl {

template <typename T>
int gemm_launch(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                T* C, int64_t ldc, unsigned int* absmax_bits, hipStream_t stream);

namespace {

constexpr int NB = 128;   // diagonal block (LDS-resident GEPP, inverted for the solves)
constexpr int NBO = 256;  // outer panel width: the trailing update runs with K = NBO

// ---------------------------------------------------------------- diagonal block GEPP
// 1024 lanes (16 waves).  Step c: the pivot (first row of largest |S[r][c]|,
// r >= c) is read from 16 per-wave candidates written during step c-1, rows c
// and p are exchanged, then wave w updates rows r = c+1+w, c+17+w, ... with
// one batched round of LDS loads: l = S[r][c] / S[c][c] (LAPACK getf2 scales
// by the reciprocal), S[r][j] -= l S[c][j] for j > c, and lane 0 (column c+1)
// records its rows' candidate for step c+1.  Two barriers per column.
constexpr int DG_WAVES = 16;
constexpr int DG_ROWS = (NB + DG_WAVES - 1) / DG_WAVES;  // rows per wave per step

template <typename T>
__global__ __launch_bounds__(1024) void diag_getrf_kernel(T* A, int64_t lda, int nb, int64_t k, int32_t* ipiv,
                                                          int32_t* info)
{
    __shared__ T S[NB * (NB + 1)];
    __shared__ T cand_v[2][DG_WAVES];
    __shared__ int cand_i[2][DG_WAVES];
    constexpr int LS = NB + 1;
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    for (int idx = tid; idx < nb * nb; idx += 1024) S[(idx / nb) * LS + idx % nb] = A[(idx / nb) * lda + idx % nb];
    __syncthreads();
    // candidates for column 0
    if (lane == 0) {
        T bv = T(-1);
        int bi = 0;
        for (int r = wave; r < nb; r += DG_WAVES) {
            T v = (T)fabs((double)S[r * LS]);
            if (v > bv) {
                bv = v;
                bi = r;
            }
        }
        cand_v[0][wave] = bv;
        cand_i[0][wave] = bi;
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        const int buf = c & 1;
        if (wave < 2) {
            // lanes 0..15 hold the 16 wave candidates; first index of the max
            T bv = lane < DG_WAVES ? cand_v[buf][lane] : T(-1);
            int p = lane < DG_WAVES ? cand_i[buf][lane] : 1 << 30;
#pragma unroll
            for (int off = 8; off > 0; off /= 2) {
                const T ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(p, off);
                if (ov > bv || (ov == bv && oi < p)) {
                    bv = ov;
                    p = oi;
                }
            }
            p = __shfl(p, 0);
            bv = __shfl(bv, 0);
            if (tid == 0) {
                ipiv[k + c] = (int32_t)(k + p);
                if (bv == T(0)) atomicCAS(info, 0, (int32_t)(k + c + 1));
            }
            if (p != c && tid < nb) {
                T t = S[c * LS + tid];
                S[c * LS + tid] = S[p * LS + tid];
                S[p * LS + tid] = t;
            }
        }
        __syncthreads();
        const T pv = S[c * LS + c];
        const T rp = pv != T(0) ? T(1) / pv : T(1);
        const int j0 = c + 1 + lane, j1 = j0 + 64;
        const T u0 = j0 < nb ? S[c * LS + j0] : T(0);
        const T u1 = j1 < nb ? S[c * LS + j1] : T(0);
        T l[DG_ROWS], x0[DG_ROWS], x1[DG_ROWS];
#pragma unroll
        for (int i = 0; i < DG_ROWS; ++i) {
            const int r = c + 1 + wave + DG_WAVES * i;
            const bool ok = r < nb;
            l[i] = ok ? S[r * LS + c] * rp : T(0);
            x0[i] = (ok && j0 < nb) ? S[r * LS + j0] : T(0);
            x1[i] = (ok && j1 < nb) ? S[r * LS + j1] : T(0);
        }
        T nb_v = T(-1);
        int nb_i = 0;
#pragma unroll
        for (int i = 0; i < DG_ROWS; ++i) {
            const int r = c + 1 + wave + DG_WAVES * i;
            if (r < nb) {
                x0[i] -= l[i] * u0;
                x1[i] -= l[i] * u1;
                if (lane == 0) S[r * LS + c] = l[i];
                if (j0 < nb) S[r * LS + j0] = x0[i];
                if (j1 < nb) S[r * LS + j1] = x1[i];
                const T a = (T)fabs((double)x0[i]);
                if (a > nb_v) {
                    nb_v = a;
                    nb_i = r;
                }
            }
        }
        if (lane == 0) {
            cand_v[buf ^ 1][wave] = nb_v;
            cand_i[buf ^ 1][wave] = nb_i;
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * nb; idx += 1024) A[(idx / nb) * lda + idx % nb] = S[(idx / nb) * LS + idx % nb];
}

// ---------------------------------------------------------------- triangular inverses
// grid (2, NB / 16): blockIdx.x 0 forms X = inv(L11) (unit lower), 1 forms
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

// ---------------------------------------------------------------- triangular solves
// One block step of the forward (LOWER) or backward substitution on nrhs <= 8
// right-hand sides stored row-major with ld = nrhs, in two launches:
//   trsv_diag:   y = inv(T_bb) src[b-rows]  -> dst[b-rows]      (one workgroup)
//   trsv_update: src[r] -= T[r, b-cols] y   for r below block b (LOWER) or
//                above it (upper), 8 lanes per row, each lane 16 consecutive
//                entries of the row (coalesced), reduced across the 8 lanes.
template <typename T>
__global__ __launch_bounds__(256) void trsv_diag_kernel(const T* inv, int64_t b0, int nbk, const T* src, T* dst,
                                                        int nrhs)
{
    __shared__ T xs[NB * 8];
    __shared__ T part[2][NB * 8];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nbk * nrhs; idx += 256) xs[idx] = src[b0 * nrhs + idx];
    __syncthreads();
    const int i = tid % NB, h = tid / NB;  // row i of inv, half h of the k range
    T acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = T(0);
    if (i < nbk) {
        const T* row = inv + (int64_t)i * NB;
        const int k0 = h * (NB / 2), k1 = min(nbk, k0 + NB / 2);
#pragma unroll 8
        for (int kk = k0; kk < k1; ++kk) {
            const T a = row[kk];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < nrhs) acc[q] += a * xs[kk * nrhs + q];
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (q < nrhs) part[h][i * nrhs + q] = acc[q];
    __syncthreads();
    for (int idx = tid; idx < nbk * nrhs; idx += 256) dst[b0 * nrhs + idx] = part[0][idx] + part[1][idx];
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_update_kernel(const T* LU, int64_t lda, int64_t n, int64_t b0, int nbk,
                                                          const T* ysrc, T* src, int nrhs)
{
    __shared__ T ys[NB * 8];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nbk * nrhs; idx += 256) ys[idx] = ysrc[b0 * nrhs + idx];
    __syncthreads();
    const int64_t lo = LOWER ? b0 + nbk : 0;
    const int64_t hi = LOWER ? n : b0;
    const int sub = tid % 8;                       // 16-entry segment of the row
    const int64_t r = lo + (int64_t)blockIdx.x * 32 + tid / 8;
    T acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = T(0);
    if (r < hi) {
        const T* row = LU + r * lda + b0;
        const int k0 = sub * 16;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            if (k0 + kk < nbk) {
                const T a = row[k0 + kk];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (q < nrhs) acc[q] += a * ys[(k0 + kk) * nrhs + q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        T v = acc[q];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        acc[q] = v;
    }
    if (r < hi && sub < nrhs) {
        T v = T(0);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q == sub) v = acc[q];
        src[r * nrhs + sub] -= v;
    }
}

// dst[i][:] = scale_c * src[perm[i]][:] with a type conversion (gather of the
// right-hand side by the factor's row permutation).
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void gather_rows_kernel(const TS* src, int64_t lds, TD* dst, int64_t n, int nrhs,
                                                          const int32_t* perm, const double* inv_scale)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    const int64_t s = perm ? perm[i] : i;
    double v = (double)src[s * lds + q];
    if (inv_scale) v *= inv_scale[q];
    dst[idx] = (TD)v;
}

template <typename TS, typename TD>
__global__ __launch_bounds__(256) void copy_rows_kernel(const TS* src, TD* dst, int64_t ldd, int64_t n, int nrhs)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    dst[i * ldd + q] = (TD)src[idx];
}

// ---------------------------------------------------------------- refinement
// r = b - A x (FP64), 8 rows per wave (each x entry loaded once for the 8
// rows), lanes striding along the rows; x and r are n x nrhs (ld nrhs), b has ld ldb.
constexpr int RES_ROWS = 8;
__global__ __launch_bounds__(256) void residual_kernel(const double* A, int64_t lda, int64_t n, const double* b,
                                                       int64_t ldb, const double* x, double* r, int nrhs)
{
    const int lane = threadIdx.x % 64;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + threadIdx.x / 64) * RES_ROWS;
    if (row0 >= n) return;
    double acc[RES_ROWS][8];
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[i][q] = 0.0;
    for (int64_t c = lane; c < n; c += 64) {
        double xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) xv[q] = q < nrhs ? x[c * nrhs + q] : 0.0;
#pragma unroll
        for (int i = 0; i < RES_ROWS; ++i) {
            const double a = row0 + i < n ? A[(row0 + i) * lda + c] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc[i][q] += a * xv[q];
        }
    }
#pragma unroll
    for (int i = 0; i < RES_ROWS; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double s = acc[i][q];
#pragma unroll
            for (int off = 32; off > 0; off /= 2) s += __shfl_xor(s, off);
            acc[i][q] = s;
        }
    if (lane < RES_ROWS * nrhs) {
        const int i = lane / nrhs, q = lane % nrhs;
        double s = 0.0;
#pragma unroll
        for (int ii = 0; ii < RES_ROWS; ++ii)
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
                if (ii == i && qq == q) s = acc[ii][qq];
        if (row0 + i < n) r[(row0 + i) * nrhs + q] = b[(row0 + i) * ldb + q] - s;
    }
}

// Per column q of an n x nrhs (ld) array: out[q] = sum_i a_iq^2 and
// out[8 + q] = max_i |a_iq| (one workgroup per column, fixed order).
__global__ __launch_bounds__(256) void colnorm_kernel(const double* a, int64_t ld, int64_t n, double* out)
{
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
                                                        const double* scale)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs) return;
    const int64_t i = idx / nrhs;
    const int q = (int)(idx % nrhs);
    x[i * ldx + q] += scale[q] * (double)d[idx];
}

__global__ void scale_from_norms_kernel(const double* norms, double* scale, double* inv_scale, int nrhs)
{
    const int q = threadIdx.x;
    if (q >= nrhs) return;
    const double m = norms[8 + q];
    scale[q] = m > 0.0 ? m : 1.0;
    inv_scale[q] = m > 0.0 ? 1.0 / m : 1.0;
}

int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename T>
struct FactorAux {
    // aux layout: perm (n int32, padded to 256 B), then per panel p: invL (NB x NB), invU (NB x NB)
    static size_t perm_bytes(int64_t n) { return align_up((size_t)n * sizeof(int32_t), 256); }
    static size_t bytes(int64_t n) { return perm_bytes(n) + (size_t)div_up(n, NB) * 2 * NB * NB * sizeof(T); }
    static int32_t* perm(void* aux) { return static_cast<int32_t*>(aux); }
    static T* invL(void* aux, int64_t n, int64_t p)
    {
        return reinterpret_cast<T*>(static_cast<char*>(aux) + perm_bytes(n)) + (size_t)p * 2 * NB * NB;
    }
    static T* invU(void* aux, int64_t n, int64_t p) { return invL(aux, n, p) + NB * NB; }
};

template <typename T>
size_t getrf_ws_bytes(int64_t n)
{
    return align_up((size_t)n * NB * sizeof(T), 256) + 256;
}

template <typename T>
int getrf_impl(T* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info_host, int* fallback_panels,
               void* ws, size_t ws_bytes, hipStream_t st)
{
    FL_CHECK(n >= 0, "negative n");
    FL_CHECK(lda >= n, "lda < n");
    if (n == 0) return FL_OK;
    FL_CHECK(A && ipiv && aux, "null pointer argument");
    if (ws_bytes < getrf_ws_bytes<T>(n) || !ws)
        return set_error(FL_EWORKSPACE, "getrf needs %zu workspace bytes", getrf_ws_bytes<T>(n));
    T* backup = static_cast<T*>(ws);
    char* tail = static_cast<char*>(ws) + align_up((size_t)n * NB * sizeof(T), 256);
    unsigned int* flag = reinterpret_cast<unsigned int*>(tail);
    int32_t* info_d = reinterpret_cast<int32_t*>(tail + 64);
    FL_HIP(hipMemsetAsync(info_d, 0, sizeof(int32_t), st));
    int fallbacks = 0;
    auto swap_cols = [&](int64_t k0, int nb, int64_t c0, int64_t c1) -> int {
        if (c1 <= c0) return FL_OK;
        hipLaunchKernelGGL(laswp_kernel<T>, dim3((unsigned)div_up(c1 - c0, 256)), dim3(256), 0, st, A, lda, k0, nb,
                           (const int32_t*)ipiv, c0, c1);
        FL_LAUNCHED();
        return FL_OK;
    };
    // Outer panels of NBO columns (the trailing update has K = NBO); each is
    // factored as sub-panels of NB columns.
    for (int64_t k = 0; k < n; k += NBO) {
        const int64_t pe = std::min<int64_t>(n, k + NBO);
        for (int64_t ks = k; ks < pe; ks += NB) {
            const int nb = (int)std::min<int64_t>(NB, pe - ks);
            const int64_t m = n - ks;
            const int64_t p = ks / NB;
            T* Akk = A + ks * lda + ks;
            T* invL = FactorAux<T>::invL(aux, n, p);
            T* invU = FactorAux<T>::invU(aux, n, p);
            if (m > nb) {
                FL_HIP(hipMemcpy2DAsync(backup, nb * sizeof(T), Akk, lda * sizeof(T), nb * sizeof(T), m,
                                        hipMemcpyDeviceToDevice, st));
                FL_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), st));
            }
            hipLaunchKernelGGL(diag_getrf_kernel<T>, dim3(1), dim3(1024), 0, st, Akk, lda, nb, ks, ipiv, info_d);
            FL_LAUNCHED();
            hipLaunchKernelGGL(trtri_kernel<T>, dim3(2, NB / TR_COLS), dim3(128), 0, st, (const T*)Akk, lda, nb, invL,
                               invU);
            FL_LAUNCHED();
            if (m > nb) {
                // L21 = A21 inv(U11), in place, with max |L21| recorded
                if (int s = gemm_launch<T>(m - nb, nb, nb, T(1), Akk + nb * lda, lda, invU, NB, T(0), Akk + nb * lda,
                                           lda, flag, st))
                    return s;
                unsigned int bits = 0;
                FL_HIP(hipMemcpyAsync(&bits, flag, sizeof(bits), hipMemcpyDeviceToHost, st));
                FL_HIP(hipStreamSynchronize(st));
                float lmax;
                std::memcpy(&lmax, &bits, sizeof(float));
                if (!(lmax <= 1.0f + 1.0f / (1 << 20))) {
                    // restore the sub-panel and pivot over every row below the diagonal
                    ++fallbacks;
                    FL_HIP(hipMemcpy2DAsync(Akk, lda * sizeof(T), backup, nb * sizeof(T), nb * sizeof(T), m,
                                            hipMemcpyDeviceToDevice, st));
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
                }
            }
            // this sub-panel's interchanges on the rest of the outer panel
            if (int s = swap_cols(ks, nb, k, ks)) return s;
            if (int s = swap_cols(ks, nb, ks + nb, pe)) return s;
            // update the outer panel's remaining columns
            const int64_t rest = pe - ks - nb;
            if (rest > 0) {
                T* A12 = Akk + nb;
                if (int s = gemm_launch<T>(nb, rest, nb, T(1), invL, NB, A12, lda, T(0), A12, lda, nullptr, st))
                    return s;
                if (m > nb)
                    if (int s = gemm_launch<T>(m - nb, rest, nb, T(-1), Akk + nb * lda, lda, A12, lda, T(1),
                                               Akk + nb * lda + nb, lda, nullptr, st))
                        return s;
            }
        }
        const int nbo = (int)(pe - k);
        // the outer panel's interchanges to its left and right
        if (int s = swap_cols(k, nbo, 0, k)) return s;
        if (int s = swap_cols(k, nbo, pe, n)) return s;
        const int64_t nr = n - pe;
        if (nr > 0) {
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
            // trailing update with K = NBO
            if (int s = gemm_launch<T>(nr, nr, nbo, T(-1), A + pe * lda + k, lda, A + k * lda + pe, lda, T(1),
                                       A + pe * lda + pe, lda, nullptr, st))
                return s;
        }
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

// Solves in place on a contiguous n x nrhs buffer `y` that already holds the
// permuted right-hand side; tmp is another n x nrhs buffer.  Result in y.
template <typename T>
int solve_permuted(const T* LU, int64_t n, int64_t lda, void* aux, T* y, T* tmp, int nrhs, hipStream_t st)
{
    const int64_t nblk = div_up(n, NB);
    // forward: L z = y  (z_b written to tmp; rows of y below block b updated)
    for (int64_t p = 0; p < nblk; ++p) {
        const int64_t b0 = p * NB;
        const int nbk = (int)std::min<int64_t>(NB, n - b0);
        hipLaunchKernelGGL(trsv_diag_kernel<T>, dim3(1), dim3(256), 0, st, (const T*)FactorAux<T>::invL(aux, n, p), b0,
                           nbk, (const T*)y, tmp, nrhs);
        FL_LAUNCHED();
        const int64_t rows = n - b0 - nbk;
        if (rows > 0) {
            hipLaunchKernelGGL((trsv_update_kernel<T, true>), dim3((unsigned)div_up(rows, 32)), dim3(256), 0, st, LU,
                               lda, n, b0, nbk, (const T*)tmp, y, nrhs);
            FL_LAUNCHED();
        }
    }
    // backward: U x = z  (x_b written to y; rows of tmp above block b updated)
    for (int64_t p = nblk - 1; p >= 0; --p) {
        const int64_t b0 = p * NB;
        const int nbk = (int)std::min<int64_t>(NB, n - b0);
        hipLaunchKernelGGL(trsv_diag_kernel<T>, dim3(1), dim3(256), 0, st, (const T*)FactorAux<T>::invU(aux, n, p), b0,
                           nbk, (const T*)tmp, y, nrhs);
        FL_LAUNCHED();
        if (b0 > 0) {
            hipLaunchKernelGGL((trsv_update_kernel<T, false>), dim3((unsigned)div_up(b0, 32)), dim3(256), 0, st, LU,
                               lda, n, b0, nbk, (const T*)y, tmp, nrhs);
            FL_LAUNCHED();
        }
    }
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
    if (ws_bytes < 2 * one || !ws) return set_error(FL_EWORKSPACE, "getrs needs %zu workspace bytes", 2 * one);
    void* aux = const_cast<void*>(aux_c);
    T* y = static_cast<T*>(ws);
    T* tmp = reinterpret_cast<T*>(static_cast<char*>(ws) + one);
    const unsigned g = (unsigned)div_up(n * nrhs, 256);
    hipLaunchKernelGGL((gather_rows_kernel<T, T>), dim3(g), dim3(256), 0, st, (const T*)B, ldb, y, n, (int)nrhs,
                       (const int32_t*)FactorAux<T>::perm(aux), (const double*)nullptr);
    FL_LAUNCHED();
    if (int s = solve_permuted<T>(LU, n, lda, aux, y, tmp, (int)nrhs, st)) return s;
    hipLaunchKernelGGL((copy_rows_kernel<T, T>), dim3(g), dim3(256), 0, st, (const T*)y, B, ldb, n, (int)nrhs);
    FL_LAUNCHED();
    return FL_OK;
}

size_t refine_ws_bytes(int64_t n, int64_t nrhs)
{
    return 3 * align_up((size_t)n * nrhs * sizeof(double), 256) + 2 * align_up((size_t)n * nrhs * sizeof(float), 256) +
           1024;
}

}  // namespace

int refine_impl(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux_c, int64_t n,
                const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
                double* resid_out, void* ws, size_t ws_bytes, hipStream_t st)
{
    FL_CHECK(n >= 0 && nrhs >= 1 && nrhs <= 8, "need 1 <= nrhs <= 8");
    FL_CHECK(ldb >= nrhs, "ldb < nrhs");
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
    const unsigned g = (unsigned)div_up(n * nrhs, 256);
    const int q = (int)nrhs;

    // r = b (contiguous, ld nrhs) and ||b|| per column
    hipLaunchKernelGGL((gather_rows_kernel<double, double>), dim3(g), dim3(256), 0, st, b, ldb, r, n, q,
                       (const int32_t*)nullptr, (const double*)nullptr);
    FL_LAUNCHED();
    hipLaunchKernelGGL(colnorm_kernel, dim3(q), dim3(256), 0, st, (const double*)r, (int64_t)nrhs, n, bnorms);
    FL_LAUNCHED();
    FL_HIP(hipMemsetAsync(xc, 0, (size_t)n * nrhs * sizeof(double), st));
    std::vector<double> hb(16), hr(16);
    FL_HIP(hipMemcpyAsync(hb.data(), bnorms, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
    int it = 0;
    double worst = 0.0;
    for (;;) {
        // r holds the current residual b - A x (x = 0 on entry)
        hipLaunchKernelGGL(colnorm_kernel, dim3(q), dim3(256), 0, st, (const double*)r, (int64_t)nrhs, n, norms);
        FL_LAUNCHED();
        FL_HIP(hipMemcpyAsync(hr.data(), norms, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        worst = 0.0;
        for (int c = 0; c < q; ++c) {
            const double bn = std::sqrt(hb[c]);
            const double rel = bn > 0 ? std::sqrt(hr[c]) / bn : std::sqrt(hr[c]);
            if (resid_out) resid_out[c] = rel;
            worst = std::max(worst, rel);
        }
        if ((it > 0 && worst <= tol) || it >= max_iter) break;
        // d = LU32^{-1} (r / s) s, with s = max |r| per column so FP32 sees O(1) values
        hipLaunchKernelGGL(scale_from_norms_kernel, dim3(1), dim3(8), 0, st, (const double*)norms, scale, inv_scale, q);
        hipLaunchKernelGGL((gather_rows_kernel<double, float>), dim3(g), dim3(256), 0, st, (const double*)r,
                           (int64_t)nrhs, y, n, q, (const int32_t*)FactorAux<float>::perm(aux),
                           (const double*)inv_scale);
        FL_LAUNCHED();
        if (int s = solve_permuted<float>(LU32, n, lda32, aux, y, tmp, q, st)) return s;
        hipLaunchKernelGGL(axpy_rows_kernel<float>, dim3(g), dim3(256), 0, st, xc, (int64_t)nrhs, (const float*)y, n, q,
                           (const double*)scale);
        FL_LAUNCHED();
        ++it;
        hipLaunchKernelGGL(residual_kernel, dim3((unsigned)div_up(n, 4 * RES_ROWS)), dim3(256), 0, st, A64, lda64, n, b, ldb,
                           (const double*)xc, r, q);
        FL_LAUNCHED();
    }
    hipLaunchKernelGGL((copy_rows_kernel<double, double>), dim3(g), dim3(256), 0, st, (const double*)xc, x, ldb, n, q);
    FL_LAUNCHED();
    FL_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = it;
    if (!(worst <= tol)) return set_error(FL_ENOCONV, "refinement reached %.3e after %d steps (tol %.1e)", worst, it, tol);
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
    return 2 * (precision == FL_FP64 ? getrs_ws_bytes<double>(n, nrhs) : getrs_ws_bytes<float>(n, nrhs));
}
size_t refine_workspace_bytes(int64_t n, int64_t nrhs) { return refine_ws_bytes(n, nrhs); }

int sgetrf(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, void* ws, size_t wsb,
           hipStream_t st)
{
    return getrf_impl<float>(A, n, lda, ipiv, aux, info, fb, ws, wsb, st);
}
int dgetrf(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, void* ws, size_t wsb,
           hipStream_t st)
{
    return getrf_impl<double>(A, n, lda, ipiv, aux, info, fb, ws, wsb, st);
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
