// MFMA GEMM for the dense LU: C = alpha A B + beta C, all row-major.
//
// 128 x 128 output tile per 256-lane workgroup (4 wave64 in a 2 x 2 grid, each
// wave 64 x 64).  FP32 uses v_mfma_f32_32x32x2_f32 (2 x 2 blocks of 32 x 32
// per wave, K tile 32); FP64 uses v_mfma_f64_16x16x4_f64 (4 x 4 blocks of
// 16 x 16 per wave, K tile 16).  Tiles are staged through LDS from coalesced
// 16-byte loads (scalar, bounds-checked loads on edge tiles or unaligned
// operands), with the next tile prefetched into registers while the current
// one is multiplied.
//
// FP32: the accumulators start from zero and the epilogue forms
// alpha (A B) + beta C, each term rounded once as in a reference GEMM.
// (Starting them from C would round each of the K partial sums at the scale
// of |C|: in the LU's trailing updates, whose products are small against C,
// that multiplied the FP32 factorization's backward error about seventeenfold
// at n = 6,144, 2^-24 |C| per step.)  FP64 with alpha = +-1 starts them from
// (beta / alpha) C = alpha beta C, loaded before the K loop (which hides the
// load), and the epilogue is stores only: C = alpha (A B + alpha beta C); its
// per-step rounding is 2^-53 |C|.  Any other alpha takes the FP32 scheme.
// alpha == 0 is handled as C = beta C.
//
// A workgroup reads all of K for its tile before it writes C, so the call is
// safe in place when the C tile is the A tile (C = A B with N <= 128 and
// ldc = lda on the same memory) or the B tile (M <= 128, C on B's memory).
// A product whose C overlaps B therefore always takes 128-row tiles: with
// 64-row ones two workgroups would share B's column block, and one could
// overwrite rows of B that the other has still to read.
//
// Optional epilogue: the largest |C| written is folded into *absmax_bits
// (IEEE float bits, atomic max), which the LU uses to verify its pivots.
#include "fl_internal.h"
#include "mfma.h"

#include <cstdlib>

namespace fl {
namespace {

using mfma::Mma;
constexpr int TM = 128, TN = 128, WG = 256;


template <typename T>
struct GemmArgs {
    int64_t M, N, K;
    T alpha, beta;
    const T* A;
    int64_t lda;
    const T* Bm;
    int64_t ldb;
    T* C;
    int64_t ldc;
    unsigned int* absmax_bits;
    int aligned;  // A, B base and leading dimensions allow 16-byte loads
    int tiles_m, tiles_n;
    int swizzle;  // 1: XCD-aware grouped tile order
    int lower;    // only the tiles on or below the diagonal (M == N), enumerated densely
    int ntiles;   // tiles of the product (all, or the lower ones)
    int64_t zb;   // grouped launches: problem blockIdx.z's pointers are z zb bytes on
    int tm;       // tile height: 128, or 64 for narrow products (gemm_tile's WMT = 1)
};

// This workgroup's problem of a grouped launch (blockIdx.z).
template <typename T>
__device__ __forceinline__ void to_problem(GemmArgs<T>& g)
{
    g.A = zoff(g.A, g.zb);
    g.Bm = zoff(g.Bm, g.zb);
    g.C = zoff(g.C, g.zb);
    g.absmax_bits = zoff(g.absmax_bits, g.zb);
}

// Tile of this workgroup.  The grid is 1-D; workgroups are dealt to the 8
// XCDs round-robin (id % 8), so with `swizzle` each XCD is first given a
// contiguous range of logical tiles, and logical tiles are ordered in groups
// of GROUP_M tile rows (column by column inside a group), so the workgroups
// resident on one XCD share A row blocks and B column blocks in its L2.
constexpr int NXCD = 8, GROUP_M = 8;
// Row r of the lower triangle of square tiles holds r + 1 tiles: (r, 0 .. r).
__device__ __forceinline__ void lower_tile(int id, int& tm, int& tn)
{
    int r = (int)((sqrtf(8.0f * (float)id + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= id) ++r;
    while (r * (r + 1) / 2 > id) --r;
    tm = r;
    tn = id - r * (r + 1) / 2;
}

// false for a workgroup past the last tile (the lower launch is padded to a
// multiple of the XCD count so that each XCD takes a contiguous share)
template <typename T>
__device__ __forceinline__ bool tile_of(const GemmArgs<T>& g, int id, int& tm, int& tn)
{
    if (g.lower) {
        const int nl = ((g.ntiles + NXCD - 1) / NXCD) * NXCD;
        id = (id % NXCD) * (nl / NXCD) + id / NXCD;
        if (id >= g.ntiles) return false;
        lower_tile(id, tm, tn);
        return true;
    }
    const int nt = g.tiles_m * g.tiles_n;
    if (!g.swizzle) {
        tm = id / g.tiles_n;
        tn = id % g.tiles_n;
        return true;
    }
    if (nt % NXCD == 0) id = (id % NXCD) * (nt / NXCD) + id / NXCD;
    const int per_group = GROUP_M * g.tiles_n;
    const int group = id / per_group;
    const int first_m = group * GROUP_M;
    const int gm = min(GROUP_M, g.tiles_m - first_m);
    const int in = id % per_group;
    tm = first_m + in % gm;
    tn = in / gm;
    return true;
}

// One TMW x 128 tile of C for workgroup-local tile index `id` of problem g:
// WMT = 2 waves along M (TMW = 128, each wave 64 x 64) or WMT = 1 (TMW = 64,
// the four waves side by side, each 64 x 32: twice the workgroups for a
// product too narrow to fill the GPU with 128-row tiles).
template <typename T, int WMT = 2>
__device__ __forceinline__ void gemm_tile(const GemmArgs<T>& g, int id, T* As, T* Bs)
{
    using P = Mma<T>;
    constexpr int TMW = 64 * WMT;        // tile height
    constexpr int WNT = 4 / WMT;         // waves along N
    constexpr int WCOL = TN / WNT;       // columns per wave
    constexpr int NBI = 64 / P::B, NBJ = WCOL / P::B;   // MFMA blocks per wave
    constexpr int KT = P::KT;
    constexpr int VW = P::VW;
    constexpr int LA = TMW + 1;        // padded LDS row (A is stored k-major)
    constexpr int EA = TMW * KT / WG;  // A elements per lane per tile
    constexpr int EB = TN * KT / WG;
    constexpr int VA = EA / VW;        // 16-byte loads per lane per tile
    constexpr int VB = EB / VW;
    constexpr int AV_ROW = KT / VW;    // vectors per A tile row
    constexpr int BV_ROW = TN / VW;    // vectors per B tile row

    const int tid = threadIdx.x;
    const int lane = tid % 64, wave = tid / 64;
    const int wm = wave / WNT, wn = wave % WNT;
    int tm, tn;
    if (!tile_of(g, id, tm, tn)) return;
    const int64_t m0 = (int64_t)tm * TMW, n0 = (int64_t)tn * TN;
    const bool full_mn = g.aligned && m0 + TMW <= g.M && n0 + TN <= g.N;

    typename P::acc_t acc[NBI][NBJ];
    const bool use_c = g.beta != T(0) && g.alpha != T(0);
    const bool pre_c = use_c && sizeof(T) == 8 && (g.alpha == T(1) || g.alpha == T(-1));   // (beta / alpha) = alpha beta exactly
    const bool post_c = use_c && !pre_c;   // C joins once, in the epilogue
    const T cscale = pre_c ? g.alpha * g.beta : T(0);
    // interior tiles take unpredicated accesses from one lane base pointer
    const bool full_c = m0 + TMW <= g.M && n0 + TN <= g.N;
    T* cbase = g.C + (m0 + wm * 64 + P::row_lane(lane)) * g.ldc + n0 + wn * WCOL + P::d_col(lane);
    auto coff = [&](int i, int j, int v) -> int64_t { return (int64_t)(i * P::B + P::row_acc(v)) * g.ldc + j * P::B; };
#pragma unroll
    for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int j = 0; j < NBJ; ++j) acc[i][j] = P::zero();
    if (pre_c) {
        if (full_c) {
#pragma unroll
            for (int i = 0; i < NBI; ++i)
#pragma unroll
                for (int j = 0; j < NBJ; ++j)
#pragma unroll
                    for (int v = 0; v < P::NACC; ++v) acc[i][j][v] = cscale * cbase[coff(i, j, v)];
        } else {
#pragma unroll
            for (int i = 0; i < NBI; ++i)
#pragma unroll
                for (int j = 0; j < NBJ; ++j)
#pragma unroll
                    for (int v = 0; v < P::NACC; ++v) {
                        int64_t r = m0 + wm * 64 + i * P::B + P::d_row(lane, v);
                        int64_t c = n0 + wn * WCOL + j * P::B + P::d_col(lane);
                        acc[i][j][v] = (r < g.M && c < g.N) ? cscale * g.C[r * g.ldc + c] : T(0);
                    }
        }
    }

    T ra[EA], rb[EB];
    auto load = [&](int64_t k0) {
        if (full_mn && k0 + KT <= g.K) {
#pragma unroll
            for (int e = 0; e < VA; ++e) {
                int idx = e * WG + tid;
                int r = idx / AV_ROW, c = (idx % AV_ROW) * VW;
                P::unpack(*reinterpret_cast<const typename P::vec_t*>(g.A + (m0 + r) * g.lda + k0 + c), ra + e * VW);
            }
#pragma unroll
            for (int e = 0; e < VB; ++e) {
                int idx = e * WG + tid;
                int r = idx / BV_ROW, c = (idx % BV_ROW) * VW;
                P::unpack(*reinterpret_cast<const typename P::vec_t*>(g.Bm + (k0 + r) * g.ldb + n0 + c), rb + e * VW);
            }
        } else {
#pragma unroll
            for (int e = 0; e < VA; ++e) {
                int idx = e * WG + tid;
                int r = idx / AV_ROW, c = (idx % AV_ROW) * VW;
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    int64_t gr = m0 + r, gc = k0 + c + q;
                    ra[e * VW + q] = (gr < g.M && gc < g.K) ? g.A[gr * g.lda + gc] : T(0);
                }
            }
#pragma unroll
            for (int e = 0; e < VB; ++e) {
                int idx = e * WG + tid;
                int r = idx / BV_ROW, c = (idx % BV_ROW) * VW;
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    int64_t gr = k0 + r, gc = n0 + c + q;
                    rb[e * VW + q] = (gr < g.K && gc < g.N) ? g.Bm[gr * g.ldb + gc] : T(0);
                }
            }
        }
    };

    const int64_t nkt = (g.K + KT - 1) / KT;
    if (nkt > 0) load(0);
    for (int64_t kt = 0; kt < nkt; ++kt) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < VA; ++e) {
            int idx = e * WG + tid;
            int r = idx / AV_ROW, c = (idx % AV_ROW) * VW;
#pragma unroll
            for (int q = 0; q < VW; ++q) As[(c + q) * LA + r] = ra[e * VW + q];
        }
#pragma unroll
        for (int e = 0; e < VB; ++e) {
            int idx = e * WG + tid;
            int r = idx / BV_ROW, c = (idx % BV_ROW) * VW;
            typename P::vec_t v;
            if constexpr (VW == 4) v = make_float4(rb[e * 4], rb[e * 4 + 1], rb[e * 4 + 2], rb[e * 4 + 3]);
            else v = make_double2(rb[e * 2], rb[e * 2 + 1]);
            *reinterpret_cast<typename P::vec_t*>(&Bs[r * TN + c]) = v;
        }
        __syncthreads();
        if (kt + 1 < nkt) load((kt + 1) * KT);
#pragma unroll
        for (int kk = 0; kk < KT; kk += P::KS) {
            const int k = kk + P::a_k(lane);
            T a[NBI], b[NBJ];
#pragma unroll
            for (int i = 0; i < NBI; ++i) a[i] = As[k * LA + wm * 64 + i * P::B + P::a_row(lane)];
#pragma unroll
            for (int j = 0; j < NBJ; ++j) b[j] = Bs[k * TN + wn * WCOL + j * P::B + P::a_row(lane)];
#pragma unroll
            for (int i = 0; i < NBI; ++i)
#pragma unroll
                for (int j = 0; j < NBJ; ++j) acc[i][j] = P::mma(a[i], b[j], acc[i][j]);
        }
    }

    const T outscale = g.alpha != T(0) ? g.alpha : T(1);
    float local_max = 0.0f;
    if (full_c) {
        // one row of blocks at a time, its loads of C issued before its stores
        // (with a run-time ldc the compiler cannot tell C's entries apart and
        // would otherwise wait for each load before the store after it)
#pragma unroll
        for (int i = 0; i < NBI; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            T cv[NBJ][P::NACC];
            if (post_c)
#pragma unroll
                for (int j = 0; j < NBJ; ++j)
#pragma unroll
                    for (int v = 0; v < P::NACC; ++v) cv[j][v] = cbase[coff(i, j, v)];
#pragma unroll
            for (int j = 0; j < NBJ; ++j)
#pragma unroll
                for (int v = 0; v < P::NACC; ++v) {
                    acc[i][j][v] = post_c ? outscale * acc[i][j][v] + g.beta * cv[j][v] : outscale * acc[i][j][v];
                    cbase[coff(i, j, v)] = acc[i][j][v];
                }
        }
        if (g.absmax_bits) {
#pragma unroll
            for (int i = 0; i < NBI; ++i)
#pragma unroll
                for (int j = 0; j < NBJ; ++j)
#pragma unroll
                    for (int v = 0; v < P::NACC; ++v) local_max = fmaxf(local_max, (float)fabs((double)acc[i][j][v]));
        }
    } else {
#pragma unroll
        for (int i = 0; i < NBI; ++i)
#pragma unroll
            for (int j = 0; j < NBJ; ++j)
#pragma unroll
                for (int v = 0; v < P::NACC; ++v) {
                    int64_t r = m0 + wm * 64 + i * P::B + P::d_row(lane, v);
                    int64_t c = n0 + wn * WCOL + j * P::B + P::d_col(lane);
                    if (r < g.M && c < g.N) {
                        T val = outscale * acc[i][j][v];
                        if (post_c) val += g.beta * g.C[r * g.ldc + c];
                        g.C[r * g.ldc + c] = val;
                        local_max = fmaxf(local_max, (float)fabs((double)val));
                    }
                }
    }
    if (g.absmax_bits) {
#pragma unroll
        for (int off = 32; off > 0; off /= 2) local_max = fmaxf(local_max, __shfl_xor(local_max, off));
        if (lane == 0) atomicMax(g.absmax_bits, __float_as_uint(local_max));
    }
}

// LDS of a tile (either height; As k-major with a padded row, Bs 16-byte aligned)
#define GEMM_TILE_LDS(T)                                                                  \
    __shared__ T As[Mma<T>::KT * (TM + 1)];                                               \
    __shared__ __attribute__((aligned(16))) T Bs[Mma<T>::KT * TN]

template <typename T>
__global__ __launch_bounds__(WG, 2) void gemm_kernel(GemmArgs<T> g)
{
    GEMM_TILE_LDS(T);
    to_problem(g);
    if (g.tm == 64) gemm_tile<T, 1>(g, (int)blockIdx.x, As, Bs);
    else gemm_tile<T, 2>(g, (int)blockIdx.x, As, Bs);
}

// Two independent products in one launch (a grouped launch): workgroups
// [0, tiles1) take problem g1, the rest problem g2.  The LU pairs its
// panel's L21 = A21 inv(U11) with the in-panel U12 = inv(L11) A12, whose
// outputs are disjoint, to save a launch per sub-panel.
template <typename T>
__global__ __launch_bounds__(WG, 2) void gemm_pair_kernel(GemmArgs<T> g1, GemmArgs<T> g2, int tiles1)
{
    GEMM_TILE_LDS(T);
    const int id = (int)blockIdx.x;
    to_problem(g1);
    to_problem(g2);
    if (id < tiles1) {
        if (g1.tm == 64) gemm_tile<T, 1>(g1, id, As, Bs);
        else gemm_tile<T, 2>(g1, id, As, Bs);
    } else {
        if (g2.tm == 64) gemm_tile<T, 1>(g2, id - tiles1, As, Bs);
        else gemm_tile<T, 2>(g2, id - tiles1, As, Bs);
    }
}

// C = beta C (the alpha == 0 case)
template <typename T>
__global__ __launch_bounds__(WG) void scale_kernel(T* C, int64_t ldc, int64_t M, int64_t N, T beta)
{
    const int64_t idx = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (idx >= M * N) return;
    T* p = C + (idx / N) * ldc + idx % N;
    *p = beta == T(0) ? T(0) : beta * *p;
}

// FL_GEMM_NARROW=0 keeps 128-row tiles for every product (a measurement switch).
bool narrow_tiles()
{
    static const bool on = [] {
        const char* e = std::getenv("FL_GEMM_NARROW");
        return !(e && *e == '0');
    }();
    return on;
}

template <typename T>
GemmArgs<T> gemm_args(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                      T beta, T* C, int64_t ldc, unsigned int* absmax_bits)
{
    const int vw = 16 / (int)sizeof(T);
    const int aligned = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                        (lda % vw == 0) && (ldb % vw == 0);
    static const char* sw_env = std::getenv("FL_GEMM_SWIZZLE");
    const int swizzle = sw_env && *sw_env ? std::atoi(sw_env) : 1;
    // 64-row tiles for a single product whose 128-row tiles would leave most
    // CUs idle (the panel products of a large factorization); grouped launches
    // of many small problems keep 128-row tiles, which measured faster there
    const int64_t t128 = ((M + TM - 1) / TM) * ((N + TN - 1) / TN);
    const char* bb = reinterpret_cast<const char*>(B);
    const char* cb = reinterpret_cast<const char*>(C);
    const bool c_on_b = cb < bb + ((K > 0 ? K - 1 : 0) * ldb + N) * (int64_t)sizeof(T) &&
                        bb < cb + ((M - 1) * ldc + N) * (int64_t)sizeof(T);
    const int tm = narrow_tiles() && batch().count == 1 && !c_on_b && t128 < 2 * device_cus() && M > 64 ? 64 : TM;
    const int tiles_m = (int)((M + tm - 1) / tm), tiles_n = (int)((N + TN - 1) / TN);
    return GemmArgs<T>{M,     N,          K,           alpha,   beta,    A,       lda,   B,
                       ldb,   C,          ldc,         absmax_bits, aligned, tiles_m, tiles_n, swizzle,
                       0,     tiles_m * tiles_n, batch().zb, tm};
}

}  // namespace

// Two products C_i = alpha_i A_i B_i + beta_i C_i in one launch; the
// outputs must not overlap each other's inputs or outputs (each may be in
// place on its own operand as gemm_launch allows).  alpha_i != 0.
template <typename T>
int gemm_pair_launch(int64_t M1, int64_t N1, int64_t K1, T alpha1, const T* A1, int64_t lda1, const T* B1,
                     int64_t ldb1, T beta1, T* C1, int64_t ldc1, unsigned int* absmax1, int64_t M2, int64_t N2,
                     int64_t K2, T alpha2, const T* A2, int64_t lda2, const T* B2, int64_t ldb2, T beta2, T* C2,
                     int64_t ldc2, hipStream_t stream)
{
    FL_CHECK(M1 > 0 && N1 > 0 && K1 >= 0 && M2 > 0 && N2 > 0 && K2 >= 0, "grouped GEMM needs non-empty products");
    FL_CHECK(alpha1 != T(0) && alpha2 != T(0), "grouped GEMM needs alpha != 0");
    const GemmArgs<T> g1 = gemm_args<T>(M1, N1, K1, alpha1, A1, lda1, B1, ldb1, beta1, C1, ldc1, absmax1);
    const GemmArgs<T> g2 = gemm_args<T>(M2, N2, K2, alpha2, A2, lda2, B2, ldb2, beta2, C2, ldc2, nullptr);
    const int t1 = g1.tiles_m * g1.tiles_n, t2 = g2.tiles_m * g2.tiles_n;
    hipLaunchKernelGGL(gemm_pair_kernel<T>, zgrid(dim3((unsigned)(t1 + t2))), dim3(WG), 0, stream, g1, g2, t1);
    FL_LAUNCHED();
    return FL_OK;
}

template int gemm_pair_launch<float>(int64_t, int64_t, int64_t, float, const float*, int64_t, const float*, int64_t,
                                     float, float*, int64_t, unsigned int*, int64_t, int64_t, int64_t, float,
                                     const float*, int64_t, const float*, int64_t, float, float*, int64_t,
                                     hipStream_t);
template int gemm_pair_launch<double>(int64_t, int64_t, int64_t, double, const double*, int64_t, const double*,
                                      int64_t, double, double*, int64_t, unsigned int*, int64_t, int64_t, int64_t,
                                      double, const double*, int64_t, const double*, int64_t, double, double*,
                                      int64_t, hipStream_t);

template <typename T>
int gemm_launch(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                T* C, int64_t ldc, unsigned int* absmax_bits, hipStream_t stream)
{
    FL_CHECK(M >= 0 && N >= 0 && K >= 0, "negative GEMM size");
    if (M == 0 || N == 0) return FL_OK;
    if (alpha == T(0)) {
        FL_CHECK(batch().count == 1, "a grouped GEMM needs alpha != 0");
        hipLaunchKernelGGL(scale_kernel<T>, dim3((unsigned)((M * N + WG - 1) / WG)), dim3(WG), 0, stream, C, ldc, M,
                           N, beta);
        FL_LAUNCHED();
        return FL_OK;
    }
    const GemmArgs<T> g = gemm_args<T>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, absmax_bits);
    hipLaunchKernelGGL(gemm_kernel<T>, zgrid(dim3((unsigned)(g.tiles_m * g.tiles_n))), dim3(WG), 0, stream, g);
    FL_LAUNCHED();
    return FL_OK;
}

template int gemm_launch<float>(int64_t, int64_t, int64_t, float, const float*, int64_t, const float*, int64_t, float,
                                float*, int64_t, unsigned int*, hipStream_t);
template int gemm_launch<double>(int64_t, int64_t, int64_t, double, const double*, int64_t, const double*, int64_t,
                                 double, double*, int64_t, unsigned int*, hipStream_t);

// C -= A B on the tiles of C (n x n) on or below the diagonal only (the tiles
// straddling it are updated whole): the trailing update of a symmetric
// factorization, half the flops of the full product.
template <typename T>
int gemm_lower_update(int64_t n, int64_t K, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc,
                      hipStream_t stream)
{
    FL_CHECK(n >= 0 && K >= 0, "negative GEMM size");
    if (n == 0) return FL_OK;
    GemmArgs<T> g = gemm_args<T>(n, n, K, T(-1), A, lda, B, ldb, T(1), C, ldc, nullptr);
    g.tm = TM;   // the lower enumeration is of square tiles
    g.tiles_m = (int)((n + TM - 1) / TM);
    g.lower = 1;
    g.ntiles = g.tiles_m * (g.tiles_m + 1) / 2;
    const int launch = (g.ntiles + NXCD - 1) / NXCD * NXCD;
    hipLaunchKernelGGL(gemm_kernel<T>, zgrid(dim3((unsigned)launch)), dim3(WG), 0, stream, g);
    FL_LAUNCHED();
    return FL_OK;
}

template int gemm_lower_update<float>(int64_t, int64_t, const float*, int64_t, const float*, int64_t, float*, int64_t,
                                      hipStream_t);
template int gemm_lower_update<double>(int64_t, int64_t, const double*, int64_t, const double*, int64_t, double*,
                                       int64_t, hipStream_t);

}  // namespace fl
