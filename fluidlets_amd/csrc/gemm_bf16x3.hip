// The LU's trailing update C -= A B (FP32 operands and C, row-major) with
// split-bf16 products on gfx950's bf16 MFMA.
//
// Each FP32 operand x is split as hi = bf16(x), lo = bf16(x - hi) (round to
// nearest even; x - hi is exact in FP32, so hi + lo = x to 2^-16 relative),
// and each product is taken as a_hi b_hi + a_hi b_lo + a_lo b_hi (the dropped
// a_lo b_lo is below 2^-16 |a b|), three v_mfma_f32_32x32x16_bf16 with FP32
// sums.  gfx950 runs bf16 MFMA at 16 times its FP32 MFMA rate.
//
// Operands are split once per update into a packed layout in the order the
// MFMA reads them.  A fragment is one MFMA operand of one wave: 64 lanes x
// 8 bf16 = 1 KB, where lane l holds
//   A: row 32 rb + l % 32, columns 16 kb + 8 (l / 32) .. + 7  (32 x 16 block (rb, kb))
//   B: column 32 cb + l % 32, rows 16 kb + 8 (l / 32) .. + 7  (16 x 32 block (kb, cb))
// and the hi and lo fragments of a block are adjacent: fragment (rb, kb, h)
// at ((rb KB + kb) 2 + h) 512 bf16.  A wave then reads each operand with one
// 16-byte load per lane from 1 KB of contiguous memory, and one (hi, lo) pair
// of A and one of B feed three MFMAs.  Rows and columns are zero-padded to
// whole workgroup tiles and K to a multiple of 16, so the product kernel reads
// without bounds checks.
//
// Product kernel: four waves in a 2 x 2 grid, each wave a (32 WM) x (32 WN)
// tile of A B held in accumulators that start from zero; the epilogue stores
// C - A B, so C enters one rounding instead of one per k step (starting the
// sums from C would round every partial sum at the scale of |C|, which in the
// LU's updates is far above the products').  Each k block's fragments are loaded once per workgroup
// through a register ring into an LDS double buffer (below).  Workgroups are
// dealt to the 8 XCDs round-robin, and tiles are ordered so that each XCD's
// L2 sees a compact block of C tiles (shared A row panels and B column
// panels).
//
// Optional tile filter: only the tiles on or below the diagonal (the lower
// triangle of C, for the symmetric factorization's updates).
//
// The same product kernel also runs the update with FP32 products
// (v_mfma_f32_32x32x2_f32): a fragment then holds one FP32 MFMA operand for
// four k steps (64 lanes x 4 floats = 1 KB, lane l: row 32 rb + l % 32 (A) or
// column 32 cb + l % 32 (B), k = 8 kb + l / 32 + 2 s for s = 0 .. 3), one
// fragment per block and k block of 8.
#include "fl_internal.h"

#include <algorithm>

namespace fl {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int FE = 32;          // fragment edge: rows of an A block, columns of a B block
constexpr int FK = 16;          // k per block
constexpr int FRAG = 64 * 8;    // bf16 per fragment
constexpr int WGT = 256;        // lanes per workgroup (4 waves, 2 x 2)

__device__ inline uint16_t bf16_bits(float x)
{
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));  // inf, nan
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Pack8 {
    uint4 hi, lo;
};

__device__ inline Pack8 split8(const float (&x)[8])
{
    uint16_t h[8], l[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        h[j] = bf16_bits(x[j]);
        l[j] = bf16_bits(x[j] - bf16_float(h[j]));
    }
    Pack8 p;
    p.hi = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16, h[4] | (uint32_t)h[5] << 16,
                      h[6] | (uint32_t)h[7] << 16);
    p.lo = make_uint4(l[0] | (uint32_t)l[1] << 16, l[2] | (uint32_t)l[3] << 16, l[4] | (uint32_t)l[5] << 16,
                      l[6] | (uint32_t)l[7] << 16);
    return p;
}

// A (M x K, lda) -> packed fragments of MBp row blocks x KB k blocks; one lane
// per (fragment, lane).
__global__ __launch_bounds__(256) void pack_a_kernel(const float* A, int64_t lda, int64_t M, int K, int64_t MBp,
                                                     int KB, uint16_t* out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= MBp * KB * 64) return;
    const int l = (int)(g & 63);
    const int64_t f = g >> 6;
    const int64_t rb = f / KB;
    const int kb = (int)(f % KB);
    const int64_t row = rb * FE + (l & 31);
    const int k0 = kb * FK + (l >> 5) * 8;
    float x[8];
    if (row < M && k0 + 8 <= K && (lda & 3) == 0 && ((uintptr_t)A & 15) == 0) {
        const float4 v0 = *reinterpret_cast<const float4*>(A + row * lda + k0);
        const float4 v1 = *reinterpret_cast<const float4*>(A + row * lda + k0 + 4);
        x[0] = v0.x; x[1] = v0.y; x[2] = v0.z; x[3] = v0.w;
        x[4] = v1.x; x[5] = v1.y; x[6] = v1.z; x[7] = v1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = (row < M && k0 + j < K) ? A[row * lda + k0 + j] : 0.0f;
    }
    const Pack8 p = split8(x);
    uint4* o = reinterpret_cast<uint4*>(out + (f * 2) * FRAG) + l;
    o[0] = p.hi;
    o[64] = p.lo;
}

// B (K x N, ldb) -> packed fragments of NBp column blocks x KB k blocks.
__global__ __launch_bounds__(256) void pack_b_kernel(const float* B, int64_t ldb, int64_t N, int K, int64_t NBp,
                                                     int KB, uint16_t* out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= NBp * KB * 64) return;
    const int l = (int)(g & 63);
    const int64_t f = g >> 6;
    const int64_t cb = f / KB;
    const int kb = (int)(f % KB);
    const int64_t col = cb * FE + (l & 31);
    const int k0 = kb * FK + (l >> 5) * 8;
    float x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = (col < N && k0 + j < K) ? B[(int64_t)(k0 + j) * ldb + col] : 0.0f;
    const Pack8 p = split8(x);
    uint4* o = reinterpret_cast<uint4*>(out + (f * 2) * FRAG) + l;
    o[0] = p.hi;
    o[64] = p.lo;
}

// FP32 fragments: A (M x K, lda) -> MBp row blocks x KB k blocks of 8.
__global__ __launch_bounds__(256) void pack_a_f32_kernel(const float* A, int64_t lda, int64_t M, int K, int64_t MBp,
                                                         int KB, float4* out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= MBp * KB * 64) return;
    const int l = (int)(g & 63);
    const int64_t f = g >> 6;
    const int64_t rb = f / KB;
    const int kb = (int)(f % KB);
    const int64_t row = rb * FE + (l & 31);
    const int k0 = kb * 8 + (l >> 5);
    float x[4];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) x[s2] = (row < M && k0 + 2 * s2 < K) ? A[row * lda + k0 + 2 * s2] : 0.0f;
    out[f * 64 + l] = make_float4(x[0], x[1], x[2], x[3]);
}

// FP32 fragments: B (K x N, ldb) -> NBp column blocks x KB k blocks of 8.
__global__ __launch_bounds__(256) void pack_b_f32_kernel(const float* B, int64_t ldb, int64_t N, int K, int64_t NBp,
                                                         int KB, float4* out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= NBp * KB * 64) return;
    const int l = (int)(g & 63);
    const int64_t f = g >> 6;
    const int64_t cb = f / KB;
    const int kb = (int)(f % KB);
    const int64_t col = cb * FE + (l & 31);
    const int k0 = kb * 8 + (l >> 5);
    float x[4];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) x[s2] = (col < N && k0 + 2 * s2 < K) ? B[(int64_t)(k0 + 2 * s2) * ldb + col] : 0.0f;
    out[f * 64 + l] = make_float4(x[0], x[1], x[2], x[3]);
}

struct Bf3Args {
    const uint16_t* Apk;
    const uint16_t* Bpk;
    float* C;
    int64_t M, N, ldc;
    int KB;
    int tiles_m, tiles_n;
    int lower;       // only tiles that reach the diagonal or below
    int ntiles;      // tiles in the launch (all, or the lower ones)
    int64_t awords;  // 4-byte words of the packed A operand
};

constexpr int NXCD = 8, GROUP = 8;

// (tm, tn) of logical tile `id` among tiles_m x tiles_n: dealt to XCDs in
// contiguous ranges (workgroup ids go round-robin over the 8 XCDs), then in
// groups of GROUP tile rows, column by column inside a group.  With `lower`
// the logical tiles are those on or below the diagonal (tile row t holds
// q (t + 1) tiles, q = tile height / tile width), row by row, and the launch
// is padded to a multiple of 8 workgroups; false for a workgroup past the
// last tile.
__device__ inline bool bf3_tile(const Bf3Args& g, int id, int q, int& tm, int& tn)
{
    if (g.lower) {
        const int nl = (g.ntiles + NXCD - 1) / NXCD * NXCD;
        id = (id % NXCD) * (nl / NXCD) + id / NXCD;
        if (id >= g.ntiles) return false;
        int r = (int)((sqrtf(8.0f * (float)id / (float)q + 1.0f) - 1.0f) * 0.5f);
        while (q * (r + 1) * (r + 2) / 2 <= id) ++r;
        while (q * r * (r + 1) / 2 > id) --r;
        tm = r;
        tn = id - q * r * (r + 1) / 2;
        return true;
    }
    const int nt = g.ntiles;
    if (nt % NXCD == 0) id = (id % NXCD) * (nt / NXCD) + id / NXCD;
    const int per_group = GROUP * g.tiles_n;
    const int group = id / per_group;
    const int first_m = group * GROUP;
    const int gm = min(GROUP, g.tiles_m - first_m);
    const int in = id % per_group;
    tm = first_m + in % gm;
    tn = in / gm;
    return true;
}

__device__ inline int d_row(int l, int i) { return 8 * (i / 4) + 4 * (l / 32) + (i % 4); }

// acc = 0, written as 0 * (words of the packed A operand): every word read
// there is a finite float (an FP32 fragment, or two finite bf16 halves), and
// accumulators that start from loaded values keep the loop's register
// allocation (from a literal zero, or any value not loaded, the compiler
// copies the accumulators across the loop's back edge and spills them).  All
// workgroups read the same first 128 x 64 words, which stay in L2.
template <int WM, int WN>
__device__ __forceinline__ void load_zero(const Bf3Args& g, int lane, floatx16 (&acc)[WM][WN])
{
    const float* z = reinterpret_cast<const float*>(g.Apk);
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                acc[i][j][v] = 0.0f * z[min((int64_t)((i * WN + j) * 16 + v) * 64 + lane, g.awords - 1)];
}

// The MFMA output layout of a wave's tile: lane l, value v of each 32 x 32
// block is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
// C = C - acc over the tile (edge tiles inside C only).  One row of 32 x 32
// blocks at a time: all its loads are issued before any of its stores (the
// compiler cannot tell C's entries apart with a run-time ldc, so left alone
// it would wait for each load before the store after it), and a scheduler
// barrier keeps the next row's loads below this row's stores, which bounds
// the registers live beside the accumulators.
template <int WM, int WN>
__device__ __forceinline__ void store_sub_c(const Bf3Args& g, int64_t r0, int64_t c0, int lane, bool full,
                                            const floatx16 (&acc)[WM][WN])
{
    if (full) {
        float* cb = g.C + (r0 + 4 * (lane / 32)) * g.ldc + c0 + (lane % 32);
#pragma unroll
        for (int i = 0; i < WM; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            float cv[WN][16];
#pragma unroll
            for (int j = 0; j < WN; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) cv[j][v] = cb[(int64_t)(i * FE + 8 * (v / 4) + (v % 4)) * g.ldc + j * FE];
#pragma unroll
            for (int j = 0; j < WN; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    cb[(int64_t)(i * FE + 8 * (v / 4) + (v % 4)) * g.ldc + j * FE] = cv[j][v] - acc[i][j][v];
        }
    } else {
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
            for (int j = 0; j < WN; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int64_t r = r0 + i * FE + d_row(lane, v), c = c0 + j * FE + (lane % 32);
                    if (r < g.M && c < g.N) g.C[r * g.ldc + c] -= acc[i][j][v];
                }
    }
}

// Product kernel.  Per k block a workgroup's unique fragments (A: its
// 2 WM row blocks, B: its 2 WN column blocks, hi and lo of each) are loaded
// once, a quarter by each wave, into a two-deep register ring (the loads for
// k blocks t + 2 and t + 3 are in flight while block t is multiplied), then
// written to an LDS double buffer, where every wave reads the fragments of
// its own rows and columns (16 bytes per lane, consecutive: no bank
// conflicts).  One barrier per k block.
// Workgroup barrier for the LDS ring: this wave's LDS accesses complete, then
// s_barrier.  Unlike __syncthreads() it does not wait for the global loads
// still in flight (the register ring's next k blocks); the "memory" clobber
// keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int WM, int WN, bool F32>
__global__ __launch_bounds__(WGT, (WM * WN <= 8) ? 2 : 1) void gemm_bf16x3_lds_kernel(Bf3Args g)
{
    constexpr int FPB = F32 ? 1 : 2;         // fragments per block and k block (FP32; bf16 hi and lo)
    constexpr int SA = 2 * FPB * WM;         // A fragments per k block (2 WM row blocks)
    constexpr int SF = SA + 2 * FPB * WN;    // fragments per k block
    constexpr int NQ = SF / 4;               // loaded by each wave
    __shared__ u32x4 lds[2][SF * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / 2, wn = wave % 2;
    int tm, tn;
    if (!bf3_tile(g, (int)blockIdx.x, WM / WN, tm, tn)) return;
    const int64_t rbW = (int64_t)tm * 2 * WM, cbW = (int64_t)tn * 2 * WN;   // workgroup's first row / column block
    const int64_t r0 = (rbW + wm * WM) * FE, c0 = (cbW + wn * WN) * FE;
    const bool full = r0 + WM * FE <= g.M && c0 + WN * FE <= g.N;
    const bool live = r0 < g.M && c0 < g.N;
    const int KB = g.KB;

    // this wave's share of a stage: slots wave + 4 q
    const u32x4* src[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int sl = wave + 4 * q;
        if (sl < SA)
            src[q] = reinterpret_cast<const u32x4*>(g.Apk) + ((rbW + sl / FPB) * KB * FPB + sl % FPB) * 64 + lane;
        else
            src[q] = reinterpret_cast<const u32x4*>(g.Bpk) +
                     ((cbW + (sl - SA) / FPB) * KB * FPB + (sl - SA) % FPB) * 64 + lane;
    }
    u32x4 p0[NQ], p1[NQ];
    // the scheduler barriers keep the loads where they are issued (otherwise
    // they sink towards their use and the ring loses its depth)
    auto gload = [&](int kb, u32x4 (&p)[NQ]) {
        // past the last k block the last one is loaded again (unused): an
        // unconditional load keeps the compiler from waiting on it at a join
        const int64_t k = min(kb, KB - 1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) p[q] = src[q][k * 64 * FPB];
        __builtin_amdgcn_sched_barrier(0);
    };
    auto lwrite = [&](int buf, const u32x4 (&p)[NQ]) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) lds[buf][(wave + 4 * q) * 64 + lane] = p[q];
    };

    floatx16 acc[WM][WN];
    if (live) load_zero<WM, WN>(g, lane, acc);

    auto mul = [&](int buf) {
        if constexpr (F32) {
            const floatx4* Lf = reinterpret_cast<const floatx4*>(lds[buf]);
            floatx4 a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i) a[i] = Lf[(wm * WM + i) * 64 + lane];
#pragma unroll
            for (int j = 0; j < WN; ++j) b[j] = Lf[(SA + wn * WN + j) * 64 + lane];
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int i = 0; i < WM; ++i)
#pragma unroll
                    for (int j = 0; j < WN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i][s2], b[j][s2], acc[i][j], 0, 0, 0);
            return;
        }
        const bf16x8* Ls = reinterpret_cast<const bf16x8*>(lds[buf]);
        bf16x8 ah[WM], al[WM], bh[WN], bl[WN];
#pragma unroll
        for (int i = 0; i < WM; ++i) {
            ah[i] = Ls[((wm * WM + i) * 2) * 64 + lane];
            al[i] = Ls[((wm * WM + i) * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            bh[j] = Ls[(SA + (wn * WN + j) * 2) * 64 + lane];
            bl[j] = Ls[(SA + (wn * WN + j) * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
            for (int j = 0; j < WN; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            }
    };

    gload(0, p0);
    gload(1, p1);
    lwrite(0, p0);
    gload(2, p0);
    lds_barrier();
    for (int t = 0; t < KB; t += 2) {
        // lds[0]: block t; p1: block t + 1; p0: block t + 2 (in flight)
        mul(0);
        if (t + 1 < KB) lwrite(1, p1);
        gload(t + 3, p1);
        lds_barrier();
        if (t + 1 >= KB) break;
        // lds[1]: block t + 1; p0: block t + 2; p1: block t + 3
        mul(1);
        if (t + 2 < KB) lwrite(0, p0);
        gload(t + 4, p0);
        lds_barrier();
    }

    if (!live) return;
    store_sub_c<WM, WN>(g, r0, c0, lane, full, acc);
}

// Wave tile 128 x 64 (WM = 4, WN = 2 blocks of 32 x 32), workgroup tile
// 256 x 128.  Measured at M = N = 24,064, K = 512 on MI355X: 2.2 ms (0.35 of the
// bf16 MFMA peak with C's read and write); 128 x 128 wave tiles spill, and
// fragments read straight from global memory without the LDS ring reached
// 0.29.  Packed operands are padded to 256 rows / columns.
constexpr int BF3_WM = 4, BF3_WN = 2;
constexpr int PAD = 256;

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

// Workspace of split3_update for an M x N update with K columns.
size_t split3_workspace_bytes(int64_t M, int64_t N, int64_t K)
{
    const int64_t KB = cdiv(K, FK);
    const int64_t MBp = cdiv(M, PAD) * (PAD / FE), NBp = cdiv(N, PAD) * (PAD / FE);
    return align_up((size_t)MBp * KB * 2 * FRAG * sizeof(uint16_t), 256) +
           align_up((size_t)NBp * KB * 2 * FRAG * sizeof(uint16_t), 256);
}

// C (M x N, ldc) -= A (M x K, lda) B (K x N, ldb) with split-bf16 products;
// with lower != 0 only C's tiles on and below the diagonal are updated
// (M == N; the tiles straddling the diagonal are updated whole).
int split3_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                  int64_t ldc, int lower, void* ws, size_t ws_bytes, hipStream_t st)
{
    FL_CHECK(M >= 0 && N >= 0 && K >= 0, "negative size");
    FL_CHECK(!lower || M == N, "a lower-triangle update needs M == N");
    FL_CHECK(batch().count == 1, "split3_update has no grouped form");
    if (M == 0 || N == 0 || K == 0) return FL_OK;
    FL_CHECK(ws && ws_bytes >= split3_workspace_bytes(M, N, K), "split3_update needs %zu workspace bytes",
             split3_workspace_bytes(M, N, K));
    const int KB = (int)cdiv(K, FK);
    constexpr int wm = BF3_WM, wn = BF3_WN;
    const int64_t tiles_m = cdiv(M, 2 * wm * FE), tiles_n = cdiv(N, 2 * wn * FE);
    const int64_t MBp = cdiv(M, PAD) * (PAD / FE), NBp = cdiv(N, PAD) * (PAD / FE);
    uint16_t* Apk = static_cast<uint16_t*>(ws);
    uint16_t* Bpk = Apk + align_up((size_t)MBp * KB * 2 * FRAG * sizeof(uint16_t), 256) / sizeof(uint16_t);
    hipLaunchKernelGGL(pack_a_kernel, dim3((unsigned)cdiv(MBp * KB * 64, 256)), dim3(256), 0, st, A, lda, M, (int)K,
                       MBp, KB, Apk);
    FL_LAUNCHED();
    hipLaunchKernelGGL(pack_b_kernel, dim3((unsigned)cdiv(NBp * KB * 64, 256)), dim3(256), 0, st, B, ldb, N, (int)K,
                       NBp, KB, Bpk);
    FL_LAUNCHED();
    Bf3Args g;
    g.Apk = Apk;
    g.Bpk = Bpk;
    g.awords = MBp * KB * 2 * FRAG / 2;
    g.C = C;
    g.M = M;
    g.N = N;
    g.ldc = ldc;
    g.KB = KB;
    g.tiles_m = (int)tiles_m;
    g.tiles_n = (int)tiles_n;
    g.lower = lower;
    // lower: tile row t holds q (t + 1) tiles (q = wm / wn), the last row's
    // beyond column tiles_n - 1 left to the waves' own bounds checks
    const int q = wm / wn;
    g.ntiles = lower ? (int)(q * tiles_m * (tiles_m + 1) / 2) : (int)(tiles_m * tiles_n);
    const unsigned launch = (unsigned)(lower ? (g.ntiles + NXCD - 1) / NXCD * NXCD : g.ntiles);
    hipLaunchKernelGGL((gemm_bf16x3_lds_kernel<BF3_WM, BF3_WN, false>), dim3(launch), dim3(WGT), 0, st, g);
    FL_LAUNCHED();
    return FL_OK;
}

// Workspace of f32_update (no more than split3_update's for the same sizes).
size_t f32_workspace_bytes(int64_t M, int64_t N, int64_t K)
{
    const int64_t KB = cdiv(K, 8);
    const int64_t MBp = cdiv(M, PAD) * (PAD / FE), NBp = cdiv(N, PAD) * (PAD / FE);
    return align_up((size_t)MBp * KB * 64 * sizeof(float4), 256) + align_up((size_t)NBp * KB * 64 * sizeof(float4), 256);
}

// C (M x N, ldc) -= A (M x K, lda) B (K x N, ldb) with FP32 MFMA products on
// packed fragments (the same product kernel as split3_update); lower as there.
int f32_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
               int64_t ldc, int lower, void* ws, size_t ws_bytes, hipStream_t st)
{
    FL_CHECK(M >= 0 && N >= 0 && K >= 0, "negative size");
    FL_CHECK(!lower || M == N, "a lower-triangle update needs M == N");
    FL_CHECK(batch().count == 1, "f32_update has no grouped form");
    if (M == 0 || N == 0 || K == 0) return FL_OK;
    FL_CHECK(ws && ws_bytes >= f32_workspace_bytes(M, N, K), "f32_update needs %zu workspace bytes",
             f32_workspace_bytes(M, N, K));
    const int KB = (int)cdiv(K, 8);
    constexpr int wm = BF3_WM, wn = BF3_WN;
    const int64_t tiles_m = cdiv(M, 2 * wm * FE), tiles_n = cdiv(N, 2 * wn * FE);
    const int64_t MBp = cdiv(M, PAD) * (PAD / FE), NBp = cdiv(N, PAD) * (PAD / FE);
    float4* Apk = static_cast<float4*>(ws);
    float4* Bpk = Apk + align_up((size_t)MBp * KB * 64 * sizeof(float4), 256) / sizeof(float4);
    hipLaunchKernelGGL(pack_a_f32_kernel, dim3((unsigned)cdiv(MBp * KB * 64, 256)), dim3(256), 0, st, A, lda, M,
                       (int)K, MBp, KB, Apk);
    FL_LAUNCHED();
    hipLaunchKernelGGL(pack_b_f32_kernel, dim3((unsigned)cdiv(NBp * KB * 64, 256)), dim3(256), 0, st, B, ldb, N,
                       (int)K, NBp, KB, Bpk);
    FL_LAUNCHED();
    Bf3Args g;
    g.Apk = reinterpret_cast<const uint16_t*>(Apk);
    g.Bpk = reinterpret_cast<const uint16_t*>(Bpk);
    g.awords = MBp * KB * 64 * 4;
    g.C = C;
    g.M = M;
    g.N = N;
    g.ldc = ldc;
    g.KB = KB;
    g.tiles_m = (int)tiles_m;
    g.tiles_n = (int)tiles_n;
    g.lower = lower;
    const int q = wm / wn;
    g.ntiles = lower ? (int)(q * tiles_m * (tiles_m + 1) / 2) : (int)(tiles_m * tiles_n);
    const unsigned launch = (unsigned)(lower ? (g.ntiles + NXCD - 1) / NXCD * NXCD : g.ntiles);
    hipLaunchKernelGGL((gemm_bf16x3_lds_kernel<BF3_WM, BF3_WN, true>), dim3(launch), dim3(WGT), 0, st, g);
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace fl
