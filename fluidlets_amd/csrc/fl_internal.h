// Internal helpers shared by the translation units of libfluidlets_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <type_traits>

#include "../../include/fluidlets_hip.h"

namespace fl {

// Thread-local error slot behind fl_get_last_error().
int set_error(int status, const char* fmt, ...);
void clear_error();

#define FL_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::fl::set_error(FL_EHIP, "%s failed: %s (%s:%d)", #call,                 \
                                   hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define FL_CHECK(cond, ...)                                                                \
    do {                                                                                   \
        if (!(cond)) return ::fl::set_error(FL_EINVAL, __VA_ARGS__);                        \
    } while (0)

// Checks the launch that was just enqueued.
#define FL_LAUNCHED()                                                                      \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return ::fl::set_error(FL_EHIP, "kernel launch failed: %s (%s:%d)",             \
                                   hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Grouped (batched) calls: one launch per kernel serves `count` problems laid
// out in equal slabs of `zb` bytes (every pointer of problem z is problem
// 0's plus z zb).  The batched entry points set this for the calling thread
// for the duration of the call; outside one it is (1, 0).  Kernels on the
// batched paths take zb and offset their pointers by blockIdx.z zb; their
// grids get gridDim.z = count from zgrid().
struct Batch {
    int64_t count = 1;
    int64_t zb = 0;
};
Batch& batch();
inline dim3 zgrid(dim3 g)
{
    g.z = (unsigned)batch().count;
    return g;
}
// The offset is taken on a byte pointer, not on an integer: a pointer that
// round-trips through uintptr_t loses its address space, and every access
// through it then compiles to FLAT instructions, which count on the LDS
// wait counter too (a wait for an LDS read would also wait for every global
// load in flight).
template <class T>
__device__ __forceinline__ T* zoff(T* p, int64_t zb)
{
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    return p ? reinterpret_cast<T*>(reinterpret_cast<B*>(p) + (uint64_t)blockIdx.z * (uint64_t)zb) : p;
}
// Sets the calling thread's batch for a scope.
struct BatchScope {
    Batch saved;
    BatchScope(int64_t count, int64_t zb) : saved(batch()) { batch() = Batch{count, zb}; }
    ~BatchScope() { batch() = saved; }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int device_cus();

// Kernel-sum entry used by the C ABI and by the solvers (GMRES, refinement).
// Arrays are SoA (3, M) / (3, N); fi (nullable) holds imaginary strengths;
// u_re receives the scaled real part, u_im (nullable) the imaginary part.
// ws must hold eval_workspace_size() bytes.
size_t eval_workspace_size(int64_t M, int64_t N, int nf, int precision);
// Workspace of the symmetric path (tgt == src, M == N, one blob width, real
// strengths, FP64), 0 when that path is not used for this size; eval_launch
// takes it when ws holds at least this many bytes.
size_t eval_sym_workspace_size(int64_t N, int nf, int precision);
int eval_launch(int kind, double alpha, int64_t M, int64_t N, const double* tgt, const double* src,
                const double* eps, int64_t eps_stride, const double* f, const double* fi, double w_re, double w_im,
                double* u_re, double* u_im, int precision, void* ws, size_t ws_bytes, hipStream_t stream);

// Scale applied to the raw kernel sums: 1/(8 pi) for the Stokeslet families,
// 1/(4 pi) for the Brinkmanlet.
double kernel_scale(int kind);

// Diagonal blocks (bodies of P consecutive points) of the matrix w G, each
// 3P x 3P row-major, written to KB (N / P blocks).
int assemble_blocks_launch(int kind, double alpha, int64_t N, int P, const double* X, const double* eps,
                           int64_t eps_stride, double w, double* KB, hipStream_t stream);

// Block-Jacobi preconditioner: batched LU with partial pivoting of nblk
// m x m row-major blocks in place (ipiv, perm: nblk x m int32), and
// z = blockdiag(K_b)^{-1} v on (3, N) structure-of-arrays vectors whose block
// b holds points b P .. b P + P - 1 (m = 3P).
int block_lu_launch(int64_t m, int64_t nblk, double* KB, int32_t* ipiv, int32_t* perm, int32_t* info,
                    hipStream_t stream);
int block_apply_launch(int64_t N, int P, const double* KB, const int32_t* perm, const double* v, double* z,
                       hipStream_t stream);
// Bytes of GMRES's block-Jacobi storage for N points in bodies of P.
size_t bj_bytes(int64_t N, int P);

}  // namespace fl
