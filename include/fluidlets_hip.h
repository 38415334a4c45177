/*
 * fluidlets_hip.h -- C ABI of libfluidlets_hip.so, the MI355X (gfx950) backend
 * of the Fluidlets regularized-Stokeslet hot path.
 *
 * The reference's boundary is a Python plugin API whose hot loops are Numba
 * functions and whose dense algebra is LAPACK through scipy/numpy.  Each entry
 * point below replaces one of those calls; the file:line of the call it
 * replaces (paths relative to the reference checkout) is given with it.
 *
 * Conventions (all entry points):
 *   - Every array argument is a DEVICE pointer owned by the caller.  Nothing
 *     is allocated by the library on these paths; scratch space comes from a
 *     caller buffer whose size the matching *_workspace_size() query returns.
 *   - Point and vector arrays are structure-of-arrays: a (3, N) block holds
 *     all x, then all y, then all z (index a*N + i).  Dense matrices are
 *     row-major with a leading dimension, entry (r, c) at r*ld + c, in the
 *     reference's [point, component] order: row 3i+a, column 3j+b.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the default
 *     stream) and the call returns without waiting, except where a function
 *     says it synchronises.
 *   - Return value: FL_OK (0) or a negative fl_status.  The message of the
 *     last failure on the calling thread is returned by fl_get_last_error().
 *   - Calls on different streams may run concurrently; calls that share a
 *     workspace buffer must be ordered by the caller.
 */
#ifndef FLUIDLETS_HIP_H
#define FLUIDLETS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FL_ABI_VERSION 1

typedef enum fl_status {
    FL_OK = 0,
    FL_EINVAL = -1,     /* invalid argument (ValueError at the Python boundary) */
    FL_EHIP = -2,       /* HIP runtime error */
    FL_ENOCONV = -3,    /* iterative solver did not converge (RuntimeError) */
    FL_ESINGULAR = -4,  /* zero pivot in a factorization */
    FL_EWORKSPACE = -5  /* workspace too small */
} fl_status;

/* Kernel families: src/fluidlets/kernels/{stokeslet3d,images,brinkmanlet3d}.py */
typedef enum fl_kind {
    FL_STOKESLET3D = 0,      /* kernels/stokeslet3d.py:53-64 */
    FL_WALL_STOKESLET3D = 1, /* kernels/images.py:92-104 (free Stokeslet + no-slip images) */
    FL_BRINKMANLET3D = 2     /* kernels/brinkmanlet3d.py:107-141 */
} fl_kind;

typedef enum fl_precision {
    FL_FP64 = 0,             /* all pair arithmetic and accumulation in FP64 */
    FL_FP32C = 1             /* FP32 pair arithmetic with FP64 displacements, FP64 Brinkmanlet
                                series branch, FP32 partial sums carried in FP64 per source tile */
} fl_precision;

typedef struct fl_kernel {
    int32_t kind;            /* fl_kind */
    int32_t reserved;
    double alpha;            /* Brinkmanlet inverse screening length, >= 0; ignored otherwise */
} fl_kernel;

/* ABI version (FL_ABI_VERSION). */
int fl_version(void);

/* Copies the last error message of the calling thread into buf (NUL-terminated,
 * truncated to n bytes); returns the status of that failure (FL_OK if none). */
int fl_get_last_error(char* buf, size_t n);

/* Number of compute units of the current device (for launch sizing by callers). */
int fl_device_cus(void);

/* ------------------------------------------------------------------ kernel sums
 * u_i = (1/mu) sum_j G(x_i, y_j; eps_j) f_j  for i < M.
 * Replaces the Numba loop _apply called at src/fluidlets/core/matvec.py:48
 * (loop body core/matvec.py:10-21, wrapper core/matvec.py:24-49).
 *
 *   tgt   (3, M)   target points          src (3, N) source points
 *   eps   eps[j * eps_stride]; eps_stride = 0 broadcasts one blob width
 *   f_re  (3, N)   real part of the strengths
 *   f_im  (3, N)   imaginary part, or NULL for real strengths
 *   mu_re, mu_im   viscosity (complex viscosity when mu_im != 0)
 *   u_re  (3, M)   real part of the velocities (overwritten)
 *   u_im  (3, M)   imaginary part; required when f_im or mu_im != 0, else may be NULL
 *   precision      fl_precision
 *   workspace      device buffer of at least fl_eval_workspace_size() bytes
 *                  (may be NULL when that size is 0)
 * Sources are summed in index order inside fixed source chunks whose partial
 * sums are combined in chunk order, so results are deterministic and do not
 * depend on other work in flight.
 *
 * Symmetric evaluation: when tgt == src (the same pointer), M == N,
 * eps_stride == 0, the strengths are real and precision is FL_FP64, the
 * kernels are reciprocal (G(y, x) = G(x, y)^T) and each unordered pair's block
 * is evaluated once and applied both ways.  fl_eval takes that path when the
 * workspace holds at least fl_eval_sym_workspace_size(N, 0, FL_FP64) bytes
 * (0 means the path is not used at this size: too few tile pairs to fill
 * the device, or a workspace above 4 GiB, about N > 428,000); the partial sums are again
 * combined in a fixed order.  This is the matvec of resistance problems
 * (core/solve.py:71-73 passes the points as both targets and sources).
 */
size_t fl_eval_workspace_size(int64_t M, int64_t N, int complex_strengths, int precision);
size_t fl_eval_sym_workspace_size(int64_t N, int complex_strengths, int precision);

int fl_eval(const fl_kernel* kernel, int64_t M, int64_t N,
            const double* tgt, const double* src, const double* eps, int64_t eps_stride,
            const double* f_re, const double* f_im, double mu_re, double mu_im,
            double* u_re, double* u_im, int precision,
            void* workspace, size_t workspace_bytes, void* stream);

/* fl_eval without a workspace, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It has no counterpart in the reference.  The partial sums live
 * in a fixed device array of FL_EVAL_NOALLOC_SCRATCH_BYTES per GPU, inside the
 * library, resolved once per device; the call allocates nothing.  Same
 * arguments and results as fl_eval given a workspace of exactly
 * FL_EVAL_NOALLOC_SCRATCH_BYTES bytes (bit for bit).  A sum whose partial
 * sums need more (fl_eval_workspace_size() above the limit; not the case for
 * any M and N on a 256-CU MI355X) returns FL_EWORKSPACE.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other host threads) race, and must be ordered by the caller. */
#define FL_EVAL_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int fl_eval_noalloc(const fl_kernel* kernel, int64_t M, int64_t N,
                    const double* tgt, const double* src, const double* eps, int64_t eps_stride,
                    const double* f_re, const double* f_im, double mu_re, double mu_im,
                    double* u_re, double* u_im, int precision, void* stream);

/* ------------------------------------------------------------------ fast summation (treecode)
 * Barycentric Lagrange treecode for the kernel sums of fl_eval: no
 * counterpart in the reference, which defers fast summation
 * (DESIGN.md:18, 381); its evaluate (core/matvec.py:24-49) is the direct sum.
 * The host (fluidlets_amd.treecode) builds a source tree of clusters and, per
 * block of at most 256 sorted targets, a list of segments: runs of sorted
 * sources summed directly, or one cluster's (degree+1)^3 proxy sources.  All
 * pointers are device memory; coordinates and strengths are (3, count)
 * structure-of-arrays, FP64.
 *
 * fl_bltc_weights: the proxies of ncl clusters.  cl holds per cluster the
 * source range [begin, end) in the sorted sources, the cluster's first proxy
 * index, its first chunk and its chunk count (5 int64); box its lo x, y, z,
 * hi x, y, z (6 doubles).  chunk holds per chunk the cluster and a source
 * range [begin, end) within it (3 int64), a cluster's chunks consecutive and
 * covering its range; each chunk is summed by one workgroup into ws, which
 * holds nchunk * 3 * nf * (degree + 1)^3 doubles (nf = 2 with f_im, else 1),
 * and the chunks are then added in order (the result does not depend on
 * timing).  Writes px (3, P) (the Chebyshev points of the box) and pf (3, P)
 * (the proxy strengths; pf_im for imaginary strengths f_im, both or
 * neither).  degree 1 .. 12.
 *
 * fl_bltc_eval: u = (1/mu) sum over each block's segments of G(x, y) f, with
 * one blob width eps (the proxies carry it too).  blk holds per block the
 * first sorted target, the count (<= 256) and the segment range [s0, s1) (4
 * int64); seg per segment the kind (0: sources, 1: proxies), the first index
 * and the count (3 int64).  tidx maps a sorted target to its row of u
 * (3, M), which is written in the caller's order.  Kernel, mu and complex
 * results as fl_eval. */
int fl_bltc_weights(int64_t ncl, const int64_t* cl, const double* box, int degree,
                    int64_t nchunk, const int64_t* chunk,
                    int64_t N, const double* src, const double* f_re, const double* f_im,
                    int64_t P, double* px, double* pf_re, double* pf_im, double* ws, void* stream);
int fl_bltc_eval(const fl_kernel* kernel, int64_t M, const double* tgt, const int64_t* tidx,
                 int64_t nblk, const int64_t* blk, const int64_t* seg,
                 int64_t N, const double* src, const double* f_re, const double* f_im,
                 int64_t P, const double* px, const double* pf_re, const double* pf_im,
                 double eps, double mu_re, double mu_im, double* u_re, double* u_im, void* stream);

/* ------------------------------------------------------------------ assembly
 * K[3i+a, 3j+b] = G_ab(x_i, y_j; eps_j) / mu,  i < M, j < N.
 * Replaces _fill called at src/fluidlets/core/assemble.py:45 (loop
 * core/assemble.py:15-24, wrapper 27-46).  Any of K64 (real FP64), K32 (real
 * FP32 copy, same layout, leading dimension ld32) and Kc (complex128,
 * interleaved re/im, leading dimension ld in complex elements) may be NULL;
 * all requested outputs are written in the same pass.  K64 and K32 hold
 * G / mu_re and require mu_im == 0; Kc holds G / (mu_re + i mu_im).
 */
int fl_assemble(const fl_kernel* kernel, int64_t M, int64_t N,
                const double* tgt, const double* src, const double* eps, int64_t eps_stride,
                double mu_re, double mu_im,
                double* K64, int64_t ld, float* K32, int64_t ld32, double* Kc, void* stream);

/* Nearest-neighbour (Gallagher-Smith) assembly:
 * K[3i+a, 3*nearest[k]+b] = sum_k G_ab(x_i, X_k; eps_k) / mu.
 * Replaces _fill_nearest called at src/fluidlets/core/assemble.py:100 (loop
 * core/assemble.py:49-58, wrapper 68-101).  The quadrature points must be
 * grouped by force point: order[] lists quadrature indices so that group j is
 * order[start[j] .. start[j+1]-1] (start has n_force+1 entries).  Each output
 * entry is written once, as the in-order sum over its group (no atomics).
 * quad is (3, Q).  K64 real FP64 (ld) and K32 FP32 (ld32): either may be
 * NULL, not both.
 */
int fl_assemble_nearest(const fl_kernel* kernel, int64_t M, int64_t Q, int64_t n_force,
                        const double* tgt, const double* quad, const double* eps, int64_t eps_stride,
                        const int64_t* order, const int64_t* start, double mu_re,
                        double* K64, int64_t ld, float* K32, int64_t ld32, void* stream);

/* Nearest force point of each quadrature point (squared Euclidean distance,
 * lowest index on ties): nearest[k] = argmin_j |quad_k - force_j|, quad (3, Q),
 * force (3, N), nearest (Q) int64.  Replaces the scipy cKDTree query of
 * src/fluidlets/core/assemble.py:61-65 (nearest_neighbors). */
int fl_nearest(int64_t Q, int64_t N, const double* quad, const double* force, int64_t* nearest, void* stream);

/* ------------------------------------------------------------------ dense LU
 * Blocked right-looking LU with partial pivoting, P A = L U, in place
 * (row-major, n x n, leading dimension lda >= n).  Replaces the LAPACK getrf
 * inside scipy.linalg.solve at src/fluidlets/core/solve.py:66 and inside
 * numpy.linalg.solve at src/fluidlets/swimming.py:101.
 *   ipiv   n int32 (device): row r was interchanged with row ipiv[r] (0-based,
 *          LAPACK order)
 *   aux    fl_getrf_aux_size() bytes (device), written here and read by the
 *          solves: the folded row permutation and the inverted diagonal blocks
 *   info   host int: 0, or k+1 when U[k][k] is exactly zero (LAPACK convention;
 *          the factorization still completes)
 *   general_panels  host int (nullable): panels whose pivots had to be searched
 *          over all rows below the diagonal (0 when the block-restricted search
 *          already found the partial-pivoting pivots)
 * Diagonal blocks of 128 are factored in LDS; inv(L11) A12, A21 inv(U11)
 * and the trailing updates run on MFMA (v_mfma_f32_32x32x2_f32 for
 * fl_sgetrf, v_mfma_f64_16x16x4_f64 for fl_dgetrf).  Synchronises on stream.
 */
size_t fl_getrf_aux_size(int64_t n, int precision);
size_t fl_getrf_workspace_size(int64_t n, int precision);
int fl_sgetrf(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* general_panels,
              void* workspace, size_t workspace_bytes, void* stream);
int fl_dgetrf(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* general_panels,
              void* workspace, size_t workspace_bytes, void* stream);

/* Asynchronous form of fl_?getrf for callers that keep the matrix's source
 * (the FP64 matrix, or the assembly that made it): it only enqueues work and
 * never waits on the stream.  Every diagonal block is factored without
 * interchanges and the largest multiplier of the whole matrix is checked on
 * the device.  When the stream has run the call, *status (one device word)
 * is 0 if no multiplier exceeded 1 + 2^-20 in magnitude.  Every pivot is then
 * within a factor 1 + 2^-20 of the largest entry left in its column: the
 * factorization is partial pivoting's except at near-ties (an entry below
 * the diagonal within 2^-20 relative of the pivot), where LAPACK could have
 * interchanged rows and this path does not; growth stays bounded as with
 * partial pivoting.  When no multiplier exceeds 1, A, ipiv and aux are
 * exactly what fl_?getrf writes.
 * Otherwise *status is 1 and A holds unusable values: refill it from its
 * source and call fl_?getrf.  Same workspace as fl_?getrf.  Replaces, like
 * fl_?getrf, the LU inside scipy.linalg.solve (src/fluidlets/core/solve.py:66)
 * and numpy.linalg.solve (src/fluidlets/swimming.py:101) for matrices whose
 * partial pivoting makes no interchanges, as the kernel matrices here do.
 */
int fl_sgetrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                    void* workspace, size_t workspace_bytes, void* stream);
int fl_dgetrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Symmetric factorization A = L D L^T without interchanges, for the kernel
 * matrices with one blob width (reciprocity makes them symmetric: G(y, x) =
 * G(x, y)^T), asynchronous like fl_?getrf_async.  Only A's lower triangle
 * (and the diagonal blocks) is read.  The factors are written in the format
 * of fl_?getrf_async (unit L below the diagonal, U = D L^T on and above it,
 * ipiv the identity, the same aux), so fl_?getrs, fl_refine and the other
 * consumers take them unchanged.  The trailing updates cover only the lower
 * triangle, half the flops of the LU.  *status as fl_?getrf_async (0: every
 * multiplier within 1 + 2^-20; else refill A and use fl_?getrf).  products
 * as fl_sgetrf_ex.  Same workspace as fl_?getrf.  Replaces the dense solve
 * of src/fluidlets/core/solve.py:66 (scipy.linalg.solve, whose
 * assume_a="sym" / "pos" forms are LAPACK's symmetric factorizations) for
 * these matrices. */
int fl_ssytrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                    void* workspace, size_t workspace_bytes, void* stream);
int fl_dsytrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Grouped small solves: one launch per kernel serves `count` problems of the
 * same order.  Problem z's buffers (every pointer argument, its workspace
 * included) are problem 0's plus z * slab_bytes (a multiple of 256): each
 * problem lives in its own slab of one allocation.  Each problem's results are
 * bit for bit those of the single call on its slab (the same kernels; each
 * problem is computed by its own workgroups, blockIdx.z).  No counterpart in
 * the reference, which solves one system per scipy.linalg.solve call
 * (src/fluidlets/core/solve.py:66); they batch the many small resistance
 * solves a server receives (the six-mode pattern of
 * examples/translating_sphere/run.py:45-46).
 *   fl_pad_solve_group     fl_pad_solve on each slab
 *   fl_sgetrf_async_group  fl_sgetrf_async (FP32 products) on each slab;
 *                          status per problem
 *   fl_refine_steps_group  fl_refine_steps on each slab */
int fl_pad_solve_group(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                       int64_t nrhs, int64_t count, int64_t slab_bytes, void* stream);
int fl_sgetrf_async_group(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                          int64_t count, int64_t slab_bytes, void* workspace, size_t workspace_bytes, void* stream);
int fl_refine_steps_group(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32,
                          int64_t n, const double* b, double* x, int64_t nrhs, int64_t ldb, int steps,
                          double* resid_dev, int64_t count, int64_t slab_bytes, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Explicit inverse from the factors of fl_sgetrf: Ainv = A^{-1} =
 * U^{-1} L^{-1} P (n x n row-major, ldi >= n), by blocked substitution on the
 * permuted identity with the stored inverted diagonal blocks and the LU's
 * update GEMMs (2 n^3 flop).  For callers that refine many right-hand sides
 * against one factorization: fl_refine_inv then corrects with one pass over
 * Ainv instead of two chained triangular sweeps.  No counterpart in the
 * reference (numpy.linalg.solve at src/fluidlets/swimming.py:101 refactors
 * every call); used for the swimmer's body-frame factor reuse. */
int fl_sgetri(const float* LU, int64_t n, int64_t lda, const void* aux, float* Ainv, int64_t ldi, void* stream);

/* fl_refine (frame == NULL) or fl_refine_frame with the corrections taken
 * from an explicit FP32 inverse (fl_sgetri) of the operator:
 *   x <- x + s Ainv32 (r / s)      (rotated by the frame when given).
 * Same arguments, workspace (fl_refine_workspace_size), stopping rule and
 * results as fl_refine.  Replaces, like fl_refine, the dense solve of
 * src/fluidlets/swimming.py:101 (numpy.linalg.solve). */
int fl_refine_inv(const double* A64, int64_t lda64, const float* Ainv32, int64_t ldi, int64_t n, const double* frame,
                  const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
                  double* resid_out, void* workspace, size_t workspace_bytes, void* stream);

/* Products of the FP32 LU's large trailing updates (A22 -= L21 U12 with both
 * dimensions >= 2048), for fl_sgetrf_ex and fl_sgetrf_async_ex.
 *   FL_LU_FP32    FP32 MFMA products, FP32 sums (what fl_sgetrf does).
 *   FL_LU_BF16X3  split bf16: every FP32 operand x is written as hi + lo
 *                 (hi = bf16(x), lo = bf16(x - hi), x = hi + lo to 2^-16
 *                 relative) and a b is taken as a_hi b_hi + a_hi b_lo + a_lo b_hi
 *                 on bf16 MFMA with FP32 sums: about twice as fast on gfx950,
 *                 with products good to about 2^-16 relative, for factors
 *                 that precondition FP64 iterative refinement (fl_refine). */
typedef enum fl_lu_products {
    FL_LU_FP32 = 0,
    FL_LU_BF16X3 = 1
} fl_lu_products;

/* fl_sgetrf and fl_sgetrf_async with a choice of trailing-update products
 * (fl_lu_products).  Same arguments and workspace (fl_getrf_workspace_size
 * with FL_FP32C covers both); FL_LU_FP32 gives exactly fl_sgetrf /
 * fl_sgetrf_async.  Replace, like them, the LU inside scipy.linalg.solve
 * (src/fluidlets/core/solve.py:66) and numpy.linalg.solve
 * (src/fluidlets/swimming.py:101), here as the factorization that the
 * mixed-precision solves refine. */
int fl_sgetrf_ex(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* general_panels,
                 int products, void* workspace, size_t workspace_bytes, void* stream);
int fl_sgetrf_async_ex(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                       int products, void* workspace, size_t workspace_bytes, void* stream);

/* C = alpha A B + beta C, row-major (M x K, K x N, M x N), on MFMA with FP32
 * (fl_sgemm) or FP64 (fl_dgemm) accumulation; C is not read when beta == 0.
 * alpha (A B) and beta C are each formed once and added, as in a reference
 * GEMM (for alpha = +-1 the sum starts from alpha beta C, exact as well).
 * C may be written over A (N <= 128, ldc = lda, same memory) or over B
 * (M <= 128, C on B's memory); no other overlap is allowed.
 * The LU's update kernel.  It has no counterpart call in the reference (there
 * it is the gemm inside LAPACK getrf); exported for tests and for callers
 * composing their own blocked algorithms. */
int fl_sgemm(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
             float beta, float* C, int64_t ldc, void* stream);
int fl_dgemm(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, void* stream);

/* Solves A X = B with the factors of fl_?getrf; B is n x nrhs row-major
 * (leading dimension ldb, nrhs <= 8), overwritten by X.  Replaces the getrs
 * inside scipy.linalg.solve (src/fluidlets/core/solve.py:66). */
size_t fl_getrs_workspace_size(int64_t n, int64_t nrhs, int precision);
int fl_sgetrs(const float* LU, int64_t n, int64_t lda, const void* aux, float* B, int64_t nrhs, int64_t ldb,
              void* workspace, size_t workspace_bytes, void* stream);
int fl_dgetrs(const double* LU, int64_t n, int64_t lda, const void* aux, double* B, int64_t nrhs, int64_t ldb,
              void* workspace, size_t workspace_bytes, void* stream);

/* Mixed-precision solve of A64 x = b: FP32 factors (fl_sgetrf of the FP32
 * copy of A64) with FP64 iterative refinement,
 *   x <- x + s LU32^{-1} (r / s),  r = b - A64 x in FP64,  s = max |r| per column,
 * until every column's ||r||_2 / ||b||_2 <= tol, or max_iter corrections.
 * b and x are n x nrhs row-major FP64 (ldb, nrhs <= 8).  iters_out (host)
 * receives the number of corrections, resid_out (host, nrhs doubles) the final
 * relative residuals.  Returns FL_ENOCONV if tol was not reached.
 * Replaces scipy.linalg.solve at src/fluidlets/core/solve.py:66 for FP64
 * results.  Synchronises on stream.
 */
size_t fl_refine_workspace_size(int64_t n, int64_t nrhs);
int fl_refine(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32, int64_t n,
              const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
              double* resid_out, void* workspace, size_t workspace_bytes, void* stream);

/* fl_refine with a fixed number of corrections, for solves captured in a HIP
 * graph: exactly `steps` corrections, no host synchronization and no memset
 * (only kernel launches on stream), and the final relative residual of each
 * column written to resid_dev (device, nrhs doubles; NaN if a triangular
 * solve's block wait timed out).  For the same factors and steps, x is bit
 * for bit fl_refine's when fl_refine stops after that many corrections.
 * Same workspace as fl_refine.  No counterpart in the reference (its solves
 * are scipy.linalg.solve at src/fluidlets/core/solve.py:66); used by the
 * served small solves. */
int fl_refine_steps(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32,
                    int64_t n, const double* b, double* x, int64_t nrhs, int64_t ldb, int steps, double* resid_dev,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Identity padding for a solve of order n <= n_pad captured in a HIP graph:
 * rows and columns n..n_pad-1 of K64 and K32 (row-major, leading dimension
 * n_pad; either may be null) are set to the identity's, and b (n_pad x nrhs,
 * nrhs <= 8) receives
 * rhs (n x nrhs, device) over zeros; the leading n x n block is not touched.
 * n is read on the device from n_dev (one double), so one graph serves every
 * order up to n_pad.  [[K, 0], [0, I]] has K's factors next to I and the
 * padded solution is K's over zeros.  No counterpart in the reference; used
 * by the served solves (captured small ones, and larger ones padded to a
 * multiple of 128, whose factorization is up to a quarter faster than at a
 * ragged order). */
int fl_pad_solve(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                 int64_t nrhs, void* stream);

/* fl_refine with factors of the same operator written in another frame: A64
 * acts on n = 3P point-major rows (row 3p + a is component a of point p), and
 * LU32 factors the FP32 copy of A' with A64 ~ R A' R^T, R = blockdiag of the
 * 3x3 frame (host, 9 doubles, row-major, lab = R body) over the P points.
 * Each correction is  x <- x + s R LU32^{-1} (R^T r / s).  Used by the free
 * swimmer to reuse the factors of its body-frame matrix at every time step
 * with the same body shape (src/fluidlets/swimming.py:90-101 rotates that
 * matrix into the lab frame); the residual is always taken against the exact
 * A64, so the answer does not depend on the reuse.  Workspace as fl_refine.
 */
int fl_refine_frame(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32, int64_t n,
                    const double* frame, const double* b, double* x, int64_t nrhs, int64_t ldb, double tol,
                    int max_iter, int* iters_out, double* resid_out, void* workspace, size_t workspace_bytes,
                    void* stream);

/* ------------------------------------------------------------------ GMRES
 * Restarted GMRES(restart) for  sum_j G(x_i, x_j) f_j / mu_re = U_i  with the
 * matrix-free kernel sum as the operator (real mu).  Replaces
 * scipy.sparse.linalg.gmres over the evaluate LinearOperator at
 * src/fluidlets/core/solve.py:68-79, with the same stopping rule
 * (||U - A f|| <= rtol ||U||, atol = 0), `restart` inner iterations per cycle
 * and at most `maxiter` cycles.
 *   X (3, N) points, U (3, N) right-hand side, f (3, N) initial guess on
 *   entry and solution on exit (device, SoA).
 *   inner_precision FL_FP64: plain GMRES.  FL_FP32C: flexible GMRES whose
 *   preconditioner is an inner GMRES cycle with FP32C kernel sums; outer
 *   matvecs and residuals in FP64.
 *   iters_out (host): total outer inner-iterations (one per Arnoldi step, the
 *   count the reference's pr_norm callback gives); resid_out (host): final
 *   true relative residual.  Returns FL_ENOCONV when not converged (the
 *   reference raises RuntimeError).  Synchronises on stream.
 */
size_t fl_gmres_workspace_size(int64_t N, int restart, int inner_precision);
int fl_gmres(const fl_kernel* kernel, int64_t N, const double* X, const double* eps, int64_t eps_stride,
             double mu_re, const double* U, double* f, double rtol, int restart, int maxiter,
             int inner_precision, int* iters_out, double* resid_out,
             void* workspace, size_t workspace_bytes, void* stream);

/* Block-Jacobi preconditioned GMRES for several bodies: the N points form
 * N / block_points bodies of block_points consecutive points, and
 * z = blockdiag(K_b)^{-1} v (K_b the FP64 matrix of body b, exactly the
 * diagonal block of the resistance matrix) preconditions fl_gmres's loop from
 * the right.  Same operator, stopping rule (true residual <= rtol ||U||),
 * arguments and results as fl_gmres with inner_precision FL_FP64; the
 * reference has no preconditioner (src/fluidlets/core/solve.py:68-79), so
 * this is an addition for many-body problems, which converge in far fewer
 * iterations.  3 block_points <= 512.  Returns FL_ESINGULAR when a block has
 * a zero pivot.  Synchronises on stream.
 */
size_t fl_gmres_bj_workspace_size(int64_t N, int restart, int block_points);
int fl_gmres_bj(const fl_kernel* kernel, int64_t N, const double* X, const double* eps, int64_t eps_stride,
                double mu_re, const double* U, double* f, double rtol, int restart, int maxiter, int block_points,
                int* iters_out, double* resid_out, void* workspace, size_t workspace_bytes, void* stream);

/* The preconditioner's pieces, for callers with their own Krylov loop.
 * fl_assemble_blocks: the N / block_points diagonal blocks (body b: points
 *   b P .. b P + P - 1) of the matrix of fl_assemble, each 3P x 3P row-major,
 *   one after the other in `blocks` (device).
 * fl_block_lu: partial-pivoting LU (getrf: largest |entry|, the first on
 *   ties) of nblk m x m row-major blocks in place, m <= 512; ipiv (LAPACK
 *   order, 0-based within each block) and perm (row i of the factors is row
 *   perm[i] of the block), nblk x m int32 each; info[b] = first zero pivot
 *   (1-based) of block b or 0.  Asynchronous.
 * fl_block_apply: z = blockdiag^{-1} v on (3, N) structure-of-arrays vectors
 *   with the factors of fl_block_lu (m = 3 block_points).  Asynchronous.
 */
int fl_assemble_blocks(const fl_kernel* kernel, int64_t N, int block_points, const double* X, const double* eps,
                       int64_t eps_stride, double mu_re, double* blocks, void* stream);
int fl_block_lu(int64_t m, int64_t nblk, double* blocks, int32_t* ipiv, int32_t* perm, int32_t* info, void* stream);
int fl_block_apply(int64_t N, int block_points, const double* blocks, const int32_t* perm, const double* v, double* z,
                   void* stream);

/* ------------------------------------------------------------------ free swimmer
 * The bordered system of src/fluidlets/swimming.py:90-101 (no slip, zero net
 * force, zero net torque), which the reference solves with numpy.linalg.solve
 * (swimming.py:101), solved here by a Schur complement on the factored K:
 *   [ K  C ] [f]   [v]      C = [-I | [x_n - x0]x]          (3N x 6)
 *   [ B  0 ] [w] = [0],     B = [count_n I ; [Xsum_n]x]      (6 x 3N),  w = [U; Omega]
 * fl_swim_rhs writes R = [v | C] (3N x 7 row-major, rows in [point, component]
 * order) from x (3, N) lab force points, x0_host (3 doubles, host) and
 * v (3, N) lab velocities.  The caller solves Y = K^{-1} R with fl_?getrs or
 * fl_refine (7 right-hand sides).  fl_swim_schur then forms S = B Y_C (6 x 6,
 * written to S, device) and g = B Y_v in FP64 with a fixed-order reduction,
 * solves S w = g (FP64, partial pivoting), writes w to UW (6 doubles,
 * device) and f = Y_v - Y_C w to f (3N, device).  count (N) and Xsum (3, N)
 * are the nearest-neighbour group sizes and quadrature-point sums.
 */
int fl_swim_rhs(int64_t N, const double* x, const double* x0_host, const double* v, double* R, void* stream);
int fl_swim_schur(int64_t N, const double* Y, const double* count, const double* Xsum, double* f, double* UW,
                  double* S, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDLETS_HIP_H */
