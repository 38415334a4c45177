// MFMA GEMM for the dense LU: C = beta C + alpha A B, all row-major.
//
// 128 x 128 output tile per 256-lane workgroup (4 wave64 in a 2 x 2 grid, each
// wave 64 x 64).  FP32 uses v_mfma_f32_32x32x2_f32 (2 x 2 blocks of 32 x 32
// per wave, K tile 32); FP64 uses v_mfma_f64_16x16x4_f64 (4 x 4 blocks of
// 16 x 16 per wave, K tile 16).  Tiles are staged through LDS from coalesced
// loads (consecutive lanes read consecutive columns) with the next tile
// prefetched into registers while the current one is multiplied.
//
// A workgroup reads all of K for its tile before it writes C, so the call is
// safe in place when the C tile is the A tile (C = A B with N <= 128 and
// ldc = lda on the same memory) or the B tile (M <= 128, C on B's memory).
//
// Optional epilogue: the largest |C| written is folded into *absmax_bits
// (IEEE float bits, atomic max), which the LU uses to verify its pivots.
#include "fl_internal.h"

namespace fl {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128, TN = 128, WG = 256;

template <typename T>
struct Mma;

template <>
struct Mma<float> {
    static constexpr int KT = 32;   // K tile staged in LDS
    static constexpr int B = 32;    // MFMA block edge
    static constexpr int KS = 2;    // K per instruction
    static constexpr int NB = 2;    // blocks per wave edge (64 / B)
    static constexpr int NACC = 16; // accumulator values per lane per block
    using acc_t = floatx16;
    __device__ static acc_t zero() { return acc_t{0}; }
    __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
    // Lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator i
    // holds D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
    __device__ static int a_row(int l) { return l % 32; }
    __device__ static int a_k(int l) { return l / 32; }
    __device__ static int d_row(int l, int i) { return 8 * (i / 4) + 4 * (l / 32) + (i % 4); }
    __device__ static int d_col(int l) { return l % 32; }
};

template <>
struct Mma<double> {
    static constexpr int KT = 16;
    static constexpr int B = 16;
    static constexpr int KS = 4;
    static constexpr int NB = 4;
    static constexpr int NACC = 4;
    using acc_t = doublex4;
    __device__ static acc_t zero() { return acc_t{0}; }
    __device__ static acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator i
    // holds D[4 i + l / 16][l % 16].
    __device__ static int a_row(int l) { return l % 16; }
    __device__ static int a_k(int l) { return l / 16; }
    __device__ static int d_row(int l, int i) { return 4 * i + l / 16; }
    __device__ static int d_col(int l) { return l % 16; }
};

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
};

template <typename T>
__global__ __launch_bounds__(WG) void gemm_kernel(GemmArgs<T> g)
{
    using P = Mma<T>;
    constexpr int KT = P::KT;
    constexpr int LA = TM + 1;  // padded LDS row (A is stored k-major)
    constexpr int EA = TM * KT / WG;  // A elements per lane per tile
    constexpr int EB = TN * KT / WG;
    __shared__ T As[KT * LA];
    __shared__ T Bs[KT * TN];

    const int tid = threadIdx.x;
    const int lane = tid % 64, wave = tid / 64;
    const int wm = wave / 2, wn = wave % 2;
    const int64_t m0 = (int64_t)blockIdx.y * TM, n0 = (int64_t)blockIdx.x * TN;

    typename P::acc_t acc[P::NB][P::NB];
#pragma unroll
    for (int i = 0; i < P::NB; ++i)
#pragma unroll
        for (int j = 0; j < P::NB; ++j) acc[i][j] = P::zero();

    T ra[EA], rb[EB];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int idx = e * WG + tid;
            int r = idx / KT, c = idx % KT;
            int64_t gr = m0 + r, gc = k0 + c;
            ra[e] = (gr < g.M && gc < g.K) ? g.A[gr * g.lda + gc] : T(0);
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int idx = e * WG + tid;
            int r = idx / TN, c = idx % TN;
            int64_t gr = k0 + r, gc = n0 + c;
            rb[e] = (gr < g.K && gc < g.N) ? g.Bm[gr * g.ldb + gc] : T(0);
        }
    };

    const int64_t nkt = (g.K + KT - 1) / KT;
    if (nkt > 0) load(0);
    for (int64_t kt = 0; kt < nkt; ++kt) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EA; ++e) {
            int idx = e * WG + tid;
            As[(idx % KT) * LA + idx / KT] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            int idx = e * WG + tid;
            Bs[(idx / TN) * TN + idx % TN] = rb[e];
        }
        __syncthreads();
        if (kt + 1 < nkt) load((kt + 1) * KT);
#pragma unroll
        for (int kk = 0; kk < KT; kk += P::KS) {
            const int k = kk + P::a_k(lane);
            T a[P::NB], b[P::NB];
#pragma unroll
            for (int i = 0; i < P::NB; ++i) a[i] = As[k * LA + wm * 64 + i * P::B + P::a_row(lane)];
#pragma unroll
            for (int j = 0; j < P::NB; ++j) b[j] = Bs[k * TN + wn * 64 + j * P::B + P::a_row(lane)];
#pragma unroll
            for (int i = 0; i < P::NB; ++i)
#pragma unroll
                for (int j = 0; j < P::NB; ++j) acc[i][j] = P::mma(a[i], b[j], acc[i][j]);
        }
    }

    float local_max = 0.0f;
#pragma unroll
    for (int i = 0; i < P::NB; ++i)
#pragma unroll
        for (int j = 0; j < P::NB; ++j)
#pragma unroll
            for (int v = 0; v < P::NACC; ++v) {
                int64_t r = m0 + wm * 64 + i * P::B + P::d_row(lane, v);
                int64_t c = n0 + wn * 64 + j * P::B + P::d_col(lane);
                if (r < g.M && c < g.N) {
                    T* p = g.C + r * g.ldc + c;
                    T val = g.alpha * acc[i][j][v];
                    if (g.beta != T(0)) val += g.beta * *p;
                    *p = val;
                    local_max = fmaxf(local_max, (float)fabs((double)val));
                }
            }
    if (g.absmax_bits) {
#pragma unroll
        for (int off = 32; off > 0; off /= 2) local_max = fmaxf(local_max, __shfl_xor(local_max, off));
        if (lane == 0) atomicMax(g.absmax_bits, __float_as_uint(local_max));
    }
}

}  // namespace

template <typename T>
int gemm_launch(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                T* C, int64_t ldc, unsigned int* absmax_bits, hipStream_t stream)
{
    FL_CHECK(M >= 0 && N >= 0 && K >= 0, "negative GEMM size");
    if (M == 0 || N == 0) return FL_OK;
    GemmArgs<T> g{M, N, K, alpha, beta, A, lda, B, ldb, C, ldc, absmax_bits};
    dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM));
    hipLaunchKernelGGL(gemm_kernel<T>, grid, dim3(WG), 0, stream, g);
    FL_LAUNCHED();
    return FL_OK;
}

template int gemm_launch<float>(int64_t, int64_t, int64_t, float, const float*, int64_t, const float*, int64_t, float,
                                float*, int64_t, unsigned int*, hipStream_t);
template int gemm_launch<double>(int64_t, int64_t, int64_t, double, const double*, int64_t, const double*, int64_t,
                                 double, double*, int64_t, unsigned int*, hipStream_t);

}  // namespace fl
