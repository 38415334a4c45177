// Block-Jacobi preconditioner for GMRES on several bodies: the diagonal
// blocks of the resistance matrix (one body of P consecutive points each,
// 3P x 3P, assembled by assemble_blocks_launch) are factored by partial
// pivoting and applied as z = blockdiag(K_b)^{-1} v on the (3, N) vectors of
// the kernel sums.  The reference has no preconditioner (scipy's GMRES at
// src/fluidlets/core/solve.py:68-79); this is the per-body preconditioning
// the rod carpet's 1,024 independent rods call for.
//
// block_lu_kernel: one workgroup per block, right-looking LU by panels of
// PW = 32 columns.  The panel (all rows below its top) is factored in LDS
// with partial pivoting (getf2: largest |entry|, the first on ties,
// multipliers scaled by the pivot's reciprocal), written back, its row
// interchanges applied to the other columns in order, U12 formed by
// substitution with the panel's unit lower triangle, and A22 -= L21 U12
// updated through 64-column chunks of U12 staged in LDS.  The interchanges
// are folded into a row permutation for the solves.
//
// block_apply_kernel: one workgroup per block.  The right-hand side is
// gathered by the permutation into LDS, then forward (unit lower) and
// backward (upper) substitution run by 64-row chunks: the part from the
// solved chunks is a row-by-row dot product spread over the four waves
// (coalesced row reads, fixed-order reductions), and the triangle inside the
// chunk is solved by one wave, a row per lane, with readlane broadcasts.
#include "fl_internal.h"

namespace fl {
namespace {

constexpr int BW = 256;     // workgroup
constexpr int PW = 32;      // panel width of the LU
constexpr int PPT = PW + 1; // LDS pitch of the panel
constexpr int CH = 64;      // chunk of the solves and of the staged U12
constexpr int MMAX = 512;   // largest block (3P <= 512, P <= 170)

__global__ __launch_bounds__(BW) void block_lu_kernel(double* KB, int64_t m, int32_t* ipiv_all, int32_t* perm_all,
                                                      int32_t* info_all)
{
    __shared__ double panel[MMAX * PPT];
    __shared__ double uch[PW * CH];
    __shared__ double red_v[BW / 64];
    __shared__ int red_i[BW / 64];
    __shared__ int piv_s;
    __shared__ int perm_s[MMAX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t b = blockIdx.x;
    double* A = KB + b * m * m;
    int32_t* ipiv = ipiv_all + b * m;
    int info = 0;
    for (int64_t k0 = 0; k0 < m; k0 += PW) {
        const int pw = (int)min((int64_t)PW, m - k0);
        const int rows = (int)(m - k0);
        for (int idx = tid; idx < rows * PW; idx += BW) {
            const int r = idx / PW, c = idx % PW;
            panel[r * PPT + c] = c < pw ? A[(k0 + r) * m + k0 + c] : 0.0;
        }
        __syncthreads();
        for (int c = 0; c < pw; ++c) {
            // pivot: the largest |entry| of column c at or below row c, the first on ties
            double bv = -1.0;
            int bi = rows;
            for (int r = c + tid; r < rows; r += BW) {
                const double v = fabs(panel[r * PPT + c]);
                if (v > bv) {
                    bv = v;
                    bi = r;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                red_v[wave] = bv;
                red_i[wave] = bi;
            }
            __syncthreads();
            if (tid == 0) {
                double v = red_v[0];
                int i = red_i[0];
                for (int w = 1; w < BW / 64; ++w)
                    if (red_v[w] > v || (red_v[w] == v && red_i[w] < i)) {
                        v = red_v[w];
                        i = red_i[w];
                    }
                if (i >= rows) i = c;   // no comparable entry (a NaN column): no interchange
                piv_s = i;
                ipiv[k0 + c] = (int32_t)(k0 + i);
            }
            __syncthreads();
            const int pr = piv_s;
            if (pr != c && tid < pw) {
                const double t = panel[c * PPT + tid];
                panel[c * PPT + tid] = panel[pr * PPT + tid];
                panel[pr * PPT + tid] = t;
            }
            __syncthreads();
            const double pv = panel[c * PPT + c];
            if (pv == 0.0 && info == 0) info = (int)(k0 + c + 1);
            const double rp = pv != 0.0 ? 1.0 / pv : 1.0;
            for (int r = c + 1 + tid; r < rows; r += BW) {
                const double l = panel[r * PPT + c] * rp;
                panel[r * PPT + c] = l;
                for (int j = c + 1; j < pw; ++j) panel[r * PPT + j] = fma(-l, panel[c * PPT + j], panel[r * PPT + j]);
            }
            __syncthreads();
        }
        for (int idx = tid; idx < rows * PW; idx += BW) {
            const int r = idx / PW, c = idx % PW;
            if (c < pw) A[(k0 + r) * m + k0 + c] = panel[r * PPT + c];
        }
        // the panel's interchanges on the other columns, in order
        for (int64_t col = tid; col < m; col += BW) {
            if (col >= k0 && col < k0 + pw) continue;
            for (int c = 0; c < pw; ++c) {
                const int64_t pr = ipiv[k0 + c];
                if (pr != k0 + c) {
                    const double t = A[(k0 + c) * m + col];
                    A[(k0 + c) * m + col] = A[pr * m + col];
                    A[pr * m + col] = t;
                }
            }
        }
        __syncthreads();
        const int64_t c1 = k0 + pw;
        if (c1 < m) {
            // U12 = inv(L11) A12, a column per lane
            for (int64_t col = c1 + tid; col < m; col += BW) {
                double u[PW];
#pragma unroll
                for (int r = 0; r < PW; ++r) u[r] = r < pw ? A[(k0 + r) * m + col] : 0.0;
#pragma unroll
                for (int r = 1; r < PW; ++r)
#pragma unroll
                    for (int t = 0; t < r; ++t) u[r] = fma(-panel[r * PPT + t], u[t], u[r]);
#pragma unroll
                for (int r = 0; r < PW; ++r)
                    if (r < pw) A[(k0 + r) * m + col] = u[r];
            }
            __syncthreads();
            // A22 -= L21 U12 by 64-column chunks of U12 staged in LDS
            const int r2 = rows - pw;
            for (int64_t j0 = c1; j0 < m; j0 += CH) {
                const int nc = (int)min((int64_t)CH, m - j0);
                for (int idx = tid; idx < PW * CH; idx += BW) {
                    const int t = idx / CH, jj = idx % CH;
                    uch[idx] = (t < pw && jj < nc) ? A[(k0 + t) * m + j0 + jj] : 0.0;
                }
                __syncthreads();
                for (int idx = tid; idx < r2 * nc; idx += BW) {
                    const int i = idx / nc, jj = idx % nc;
                    double* a = A + (c1 + i) * m + j0 + jj;
                    double acc = *a;
                    const double* l = panel + (pw + i) * PPT;
#pragma unroll
                    for (int t = 0; t < PW; ++t) acc = fma(-l[t], uch[t * CH + jj], acc);
                    *a = acc;
                }
                __syncthreads();
            }
        }
    }
    // the row permutation: row i of the factors is original row perm[i]
    for (int i = tid; i < m; i += BW) perm_s[i] = i;
    __syncthreads();
    if (tid == 0) {
        for (int64_t k = 0; k < m; ++k) {
            const int p = ipiv[k];
            const int t = perm_s[k];
            perm_s[k] = perm_s[p];
            perm_s[p] = t;
        }
        info_all[b] = info;
    }
    __syncthreads();
    for (int i = tid; i < m; i += BW) perm_all[b * m + i] = perm_s[i];
}

__device__ __forceinline__ double read_lane_d(double x, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                            __builtin_amdgcn_readlane(__double2loint(x), l));
}

// acc[i] (i < nc, in part[]) = sum over columns k in [k_lo, k_hi) of
// K[r0 + i][k] zs[k]: wave w takes rows i = w, w + 4, ...; lanes stride along
// the row (coalesced); each row's sum is reduced in a fixed order.
__device__ __forceinline__ void chunk_products(const double* K, int64_t m, int64_t r0, int nc, int64_t k_lo,
                                               int64_t k_hi, const double* zs, double* part, int lane, int wave)
{
    for (int i = wave; i < nc; i += BW / 64) {
        const double* row = K + (r0 + i) * m;
        double s = 0.0;
        for (int64_t k = k_lo + lane; k < k_hi; k += 64) s = fma(row[k], zs[k], s);
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) part[i] = s;
    }
}

__global__ __launch_bounds__(BW) void block_apply_kernel(int64_t N, int P, const double* KB, const int32_t* perm_all,
                                                         const double* v, double* z)
{
    __shared__ double zs[MMAX];
    __shared__ double part[CH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t b = blockIdx.x;
    const int64_t m = 3 * (int64_t)P;
    const double* K = KB + b * m * m;
    const int32_t* perm = perm_all + b * m;
    const int64_t p0 = b * P;
    for (int64_t i = tid; i < m; i += BW) {
        const int s = perm[i];
        zs[i] = v[(s % 3) * N + p0 + s / 3];
    }
    __syncthreads();
    // forward: L y = P v (unit lower)
    for (int64_t c0 = 0; c0 < m; c0 += CH) {
        const int nc = (int)min((int64_t)CH, m - c0);
        chunk_products(K, m, c0, nc, 0, c0, zs, part, lane, wave);
        __syncthreads();
        if (wave == 0) {
            const bool act = lane < nc;
            double l[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) l[k] = (act && k < lane) ? K[(c0 + lane) * m + c0 + k] : 0.0;
            double y = act ? zs[c0 + lane] - (c0 > 0 ? part[lane] : 0.0) : 0.0;
#pragma unroll
            for (int k = 0; k < CH - 1; ++k) {
                const double yk = read_lane_d(y, k);   // row k is final here
                y = fma(-l[k], yk, y);                 // l[k] = 0 for rows at or above k
            }
            if (act) zs[c0 + lane] = y;
        }
        __syncthreads();
    }
    // backward: U x = y
    const int64_t last = (m - 1) / CH * CH;
    for (int64_t c0 = last; c0 >= 0; c0 -= CH) {
        const int nc = (int)min((int64_t)CH, m - c0);
        chunk_products(K, m, c0, nc, c0 + nc, m, zs, part, lane, wave);
        __syncthreads();
        if (wave == 0) {
            const bool act = lane < nc;
            double u[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) u[k] = (act && k > lane && k < nc) ? K[(c0 + lane) * m + c0 + k] : 0.0;
            const double d = act ? K[(c0 + lane) * m + c0 + lane] : 1.0;
            double x = act ? zs[c0 + lane] - (c0 + nc < m ? part[lane] : 0.0) : 0.0;
#pragma unroll
            for (int k = CH - 1; k >= 0; --k) {
                if (lane == k) x = x / d;               // row k is final after this
                const double xk = read_lane_d(x, k);
                if (lane < k) x = fma(-u[k], xk, x);
            }
            if (act) zs[c0 + lane] = x;
        }
        __syncthreads();
    }
    for (int64_t i = tid; i < m; i += BW) z[(i % 3) * N + p0 + i / 3] = zs[i];
}

}  // namespace

int block_lu_launch(int64_t m, int64_t nblk, double* KB, int32_t* ipiv, int32_t* perm, int32_t* info,
                    hipStream_t stream)
{
    FL_CHECK(m >= 1 && m <= MMAX, "block size %lld outside 1 .. %d", (long long)m, MMAX);
    FL_CHECK(nblk >= 0, "negative block count");
    if (nblk == 0) return FL_OK;
    hipLaunchKernelGGL(block_lu_kernel, dim3((unsigned)nblk), dim3(BW), 0, stream, KB, m, ipiv, perm, info);
    FL_LAUNCHED();
    return FL_OK;
}

int block_apply_launch(int64_t N, int P, const double* KB, const int32_t* perm, const double* v, double* z,
                       hipStream_t stream)
{
    FL_CHECK(P >= 1 && 3 * (int64_t)P <= MMAX, "block of %d points outside 1 .. %d", P, MMAX / 3);
    FL_CHECK(N >= 0 && N % P == 0, "N must be a multiple of the block size");
    const int64_t nblk = N / P;
    if (nblk == 0) return FL_OK;
    hipLaunchKernelGGL(block_apply_kernel, dim3((unsigned)nblk), dim3(BW), 0, stream, N, P, KB, perm, v, z);
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace fl
