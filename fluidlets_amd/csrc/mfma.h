// MFMA operand and accumulator layouts for the wave-level matrix products of
// the GEMM and the LU's diagonal blocks (gfx950):
//   FP32: v_mfma_f32_32x32x2_f32, one 32 x 32 block, K = 2 per instruction;
//   FP64: v_mfma_f64_16x16x4_f64, one 16 x 16 block, K = 4 per instruction.
#pragma once

#include <hip/hip_runtime.h>

namespace fl {
namespace mfma {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;

template <>
struct Mma<float> {
    static constexpr int KT = 32;   // K tile staged in LDS
    static constexpr int B = 32;    // MFMA block edge
    static constexpr int KS = 2;    // K per instruction
    static constexpr int NB = 2;    // blocks per wave edge (64 / B)
    static constexpr int NACC = 16; // accumulator values per lane per block
    static constexpr int VW = 4;    // elements per 16-byte load
    using acc_t = floatx16;
    using vec_t = float4;
    __device__ static acc_t zero() { return acc_t{0}; }
    __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
    // Lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator i
    // holds D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
    __device__ static int a_row(int l) { return l % 32; }
    __device__ static int a_k(int l) { return l / 32; }
    __device__ static int d_row(int l, int i) { return 8 * (i / 4) + 4 * (l / 32) + (i % 4); }
    __device__ static int d_col(int l) { return l % 32; }
    __device__ static int row_lane(int l) { return 4 * (l / 32); }       // d_row = row_lane + row_acc
    __device__ static int row_acc(int i) { return 8 * (i / 4) + (i % 4); }
    __device__ static void unpack(const vec_t& v, float* o)
    {
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
};

template <>
struct Mma<double> {
    static constexpr int KT = 16;
    static constexpr int B = 16;
    static constexpr int KS = 4;
    static constexpr int NB = 4;
    static constexpr int NACC = 4;
    static constexpr int VW = 2;
    using acc_t = doublex4;
    using vec_t = double2;
    __device__ static acc_t zero() { return acc_t{0}; }
    __device__ static acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator i
    // holds D[4 i + l / 16][l % 16].
    __device__ static int a_row(int l) { return l % 16; }
    __device__ static int a_k(int l) { return l / 16; }
    __device__ static int d_row(int l, int i) { return 4 * i + l / 16; }
    __device__ static int d_col(int l) { return l % 16; }
    __device__ static int row_lane(int l) { return l / 16; }
    __device__ static int row_acc(int i) { return 4 * i; }
    __device__ static void unpack(const vec_t& v, double* o)
    {
        o[0] = v.x;
        o[1] = v.y;
    }
};

}  // namespace mfma
}  // namespace fl
