// extern "C" entry points of libfluidlets_hip.so: argument checks, error slot
// and dispatch to the HIP kernels.  The declarations (with the reference call
// each one replaces) are in include/fluidlets_hip.h.
#include "fl_internal.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace fl {

namespace {
thread_local int g_status = FL_OK;
thread_local char g_msg[512] = "";
}  // namespace

int set_error(int status, const char* fmt, ...)
{
    g_status = status;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error()
{
    g_status = FL_OK;
    g_msg[0] = '\0';
}

Batch& batch()
{
    static thread_local Batch b;
    return b;
}

int device_cus()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

int assemble_launch(int kind, double alpha, int64_t M, int64_t N, const double* tgt, const double* src,
                    const double* eps, int64_t eps_stride, double mu_re, double mu_im, double* K64, int64_t ld,
                    float* K32, int64_t ld32, double* Kc, hipStream_t stream);
int assemble_nearest_launch(int kind, double alpha, int64_t M, int64_t Q, int64_t n_force, const double* tgt,
                            const double* quad, const double* eps, int64_t eps_stride, const int64_t* order,
                            const int64_t* start, double mu_re, double* K64, int64_t ld, float* K32, int64_t ld32,
                            hipStream_t stream);

int nearest_launch(int64_t Q, int64_t N, const double* quad, const double* force, int64_t* out,
                   hipStream_t stream);
int bltc_weights_launch(int64_t ncl, const int64_t* cl, const double* box, int degree, int64_t nchunk,
                        const int64_t* chunk, int64_t N, const double* src, const double* f, const double* fi,
                        int64_t P, double* px, double* pf, double* pfi, double* ws, hipStream_t stream);
int bltc_eval_launch(int kind, double alpha, int64_t M, const double* tgt, const int64_t* tidx, int64_t nblk,
                     const int64_t* blk, const int64_t* seg, int64_t N, const double* src, const double* f,
                     const double* fi, int64_t P, const double* px, const double* pf, const double* pfi, double eps,
                     double w_re, double w_im, double* u_re, double* u_im, hipStream_t stream);

size_t getrf_aux_bytes(int64_t n, int precision);
size_t getrf_workspace_bytes(int64_t n, int precision);
size_t getrs_workspace_bytes(int64_t n, int64_t nrhs, int precision);
size_t refine_workspace_bytes(int64_t n, int64_t nrhs);
int sgetrf(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, int products, void* ws,
           size_t wsb, hipStream_t st);
int dgetrf(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fb, void* ws, size_t wsb,
           hipStream_t st);
int sgetrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                 void* ws, size_t wsb, hipStream_t st);
int dgetrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, void* ws,
                 size_t wsb, hipStream_t st);
int ssytrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                 void* ws, size_t wsb, hipStream_t st);
int dsytrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, void* ws,
                 size_t wsb, hipStream_t st);
int sgetrs(const float* LU, int64_t n, int64_t lda, const void* aux, float* B, int64_t nrhs, int64_t ldb, void* ws,
           size_t wsb, hipStream_t st);
int dgetrs(const double* LU, int64_t n, int64_t lda, const void* aux, double* B, int64_t nrhs, int64_t ldb, void* ws,
           size_t wsb, hipStream_t st);
int getri_impl(const float* LU, int64_t n, int64_t lda, const void* aux, float* X, int64_t ldx, hipStream_t st);
int pad_solve_impl(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                   int64_t nrhs, hipStream_t st);
int refine_impl(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux, int64_t n,
                const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
                double* resid_out, const double* frame, void* ws, size_t ws_bytes, hipStream_t st,
                const float* inv32 = nullptr, int64_t ldi = 0, double* resid_dev = nullptr);

size_t gmres_workspace_bytes(int64_t N, int restart, int inner_precision, int block_points);
int gmres_impl(int kind, double alpha, int64_t N, const double* X, const double* eps, int64_t eps_stride,
               double mu_re, const double* U, double* f, double rtol, int restart, int maxiter, int inner_precision,
               int* iters_out, double* resid_out, void* ws, size_t ws_bytes, hipStream_t st, int block_points);
size_t bj_bytes(int64_t N, int P);

template <typename T>
int gemm_launch(int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                T* C, int64_t ldc, unsigned int* absmax_bits, hipStream_t stream);

int swim_rhs_launch(int64_t N, const double* x, const double* x0, const double* v, double* R, hipStream_t st);
int swim_schur_launch(int64_t N, const double* Y, const double* count, const double* Xsum, double* f, double* UW,
                      double* S, hipStream_t st);

}  // namespace fl

using namespace fl;

namespace {

int check_kernel(const fl_kernel* k)
{
    if (!k) return set_error(FL_EINVAL, "kernel is NULL");
    if (k->kind < FL_STOKESLET3D || k->kind > FL_BRINKMANLET3D)
        return set_error(FL_EINVAL, "unknown kernel kind %d", k->kind);
    if (k->kind == FL_BRINKMANLET3D && !(k->alpha >= 0.0))
        return set_error(FL_EINVAL, "alpha must be real and non-negative");
    return FL_OK;
}

}  // namespace

extern "C" {

int fl_version(void) { return FL_ABI_VERSION; }

int fl_get_last_error(char* buf, size_t n)
{
    if (buf && n > 0) {
        std::strncpy(buf, g_msg, n - 1);
        buf[n - 1] = '\0';
    }
    return g_status;
}

int fl_device_cus(void) { return device_cus(); }

size_t fl_eval_workspace_size(int64_t M, int64_t N, int complex_strengths, int precision)
{
    return eval_workspace_size(M, N, complex_strengths ? 2 : 1, precision);
}

size_t fl_eval_sym_workspace_size(int64_t N, int complex_strengths, int precision)
{
    return eval_sym_workspace_size(N, complex_strengths ? 2 : 1, precision);
}

int fl_eval(const fl_kernel* kernel, int64_t M, int64_t N, const double* tgt, const double* src, const double* eps,
            int64_t eps_stride, const double* f_re, const double* f_im, double mu_re, double mu_im, double* u_re,
            double* u_im, int precision, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    FL_CHECK(mu_re != 0.0 || mu_im != 0.0, "mu must be nonzero");
    const double s = kernel_scale(kernel->kind);
    double w_re, w_im;
    if (mu_im == 0.0) {
        w_re = s / mu_re;
        w_im = 0.0;
    } else {
        const double den = mu_re * mu_re + mu_im * mu_im;
        w_re = s * (mu_re / den);
        w_im = -s * (mu_im / den);
    }
    return eval_launch(kernel->kind, kernel->alpha, M, N, tgt, src, eps, eps_stride, f_re, f_im, w_re, w_im, u_re,
                       u_im, precision, workspace, workspace_bytes, as_stream(stream));
}

int fl_bltc_weights(int64_t ncl, const int64_t* cl, const double* box, int degree, int64_t nchunk,
                    const int64_t* chunk, int64_t N, const double* src, const double* f_re, const double* f_im,
                    int64_t P, double* px, double* pf_re, double* pf_im, double* ws, void* stream)
{
    clear_error();
    return bltc_weights_launch(ncl, cl, box, degree, nchunk, chunk, N, src, f_re, f_im, P, px, pf_re, pf_im, ws,
                               as_stream(stream));
}

int fl_bltc_eval(const fl_kernel* kernel, int64_t M, const double* tgt, const int64_t* tidx, int64_t nblk,
                 const int64_t* blk, const int64_t* seg, int64_t N, const double* src, const double* f_re,
                 const double* f_im, int64_t P, const double* px, const double* pf_re, const double* pf_im,
                 double eps, double mu_re, double mu_im, double* u_re, double* u_im, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    FL_CHECK(mu_re != 0.0 || mu_im != 0.0, "mu must be nonzero");
    const double s = kernel_scale(kernel->kind);
    double w_re, w_im;
    if (mu_im == 0.0) {
        w_re = s / mu_re;
        w_im = 0.0;
    } else {
        const double den = mu_re * mu_re + mu_im * mu_im;
        w_re = s * (mu_re / den);
        w_im = -s * (mu_im / den);
    }
    return bltc_eval_launch(kernel->kind, kernel->alpha, M, tgt, tidx, nblk, blk, seg, N, src, f_re, f_im, P, px,
                            pf_re, pf_im, eps, w_re, w_im, u_re, u_im, as_stream(stream));
}

int fl_assemble(const fl_kernel* kernel, int64_t M, int64_t N, const double* tgt, const double* src,
                const double* eps, int64_t eps_stride, double mu_re, double mu_im, double* K64, int64_t ld,
                float* K32, int64_t ld32, double* Kc, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    return assemble_launch(kernel->kind, kernel->alpha, M, N, tgt, src, eps, eps_stride, mu_re, mu_im, K64, ld, K32,
                           ld32, Kc, as_stream(stream));
}

int fl_assemble_nearest(const fl_kernel* kernel, int64_t M, int64_t Q, int64_t n_force, const double* tgt,
                        const double* quad, const double* eps, int64_t eps_stride, const int64_t* order,
                        const int64_t* start, double mu_re, double* K64, int64_t ld, float* K32, int64_t ld32,
                        void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    return assemble_nearest_launch(kernel->kind, kernel->alpha, M, Q, n_force, tgt, quad, eps, eps_stride, order,
                                   start, mu_re, K64, ld, K32, ld32, as_stream(stream));
}

size_t fl_getrf_aux_size(int64_t n, int precision) { return getrf_aux_bytes(n, precision); }

size_t fl_getrf_workspace_size(int64_t n, int precision) { return getrf_workspace_bytes(n, precision); }

int fl_sgetrf(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fallback_panels,
              void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return sgetrf(A, n, lda, ipiv, aux, info, fallback_panels, FL_LU_FP32, workspace, workspace_bytes,
                  as_stream(stream));
}

int fl_sgetrf_ex(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fallback_panels,
                 int products, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return sgetrf(A, n, lda, ipiv, aux, info, fallback_panels, products, workspace, workspace_bytes,
                  as_stream(stream));
}

int fl_dgetrf(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, int* info, int* fallback_panels,
              void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return dgetrf(A, n, lda, ipiv, aux, info, fallback_panels, workspace, workspace_bytes, as_stream(stream));
}

int fl_sgetrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, void* workspace,
                    size_t workspace_bytes, void* stream)
{
    clear_error();
    return sgetrf_async(A, n, lda, ipiv, aux, status, FL_LU_FP32, workspace, workspace_bytes, as_stream(stream));
}

int fl_sgetrf_async_ex(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                       int products, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return sgetrf_async(A, n, lda, ipiv, aux, status, products, workspace, workspace_bytes, as_stream(stream));
}

int fl_dgetrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                    void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return dgetrf_async(A, n, lda, ipiv, aux, status, workspace, workspace_bytes, as_stream(stream));
}

int fl_ssytrf_async(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status, int products,
                    void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return ssytrf_async(A, n, lda, ipiv, aux, status, products, workspace, workspace_bytes, as_stream(stream));
}

int fl_dsytrf_async(double* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                    void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return dsytrf_async(A, n, lda, ipiv, aux, status, workspace, workspace_bytes, as_stream(stream));
}

size_t fl_getrs_workspace_size(int64_t n, int64_t nrhs, int precision)
{
    return getrs_workspace_bytes(n, nrhs, precision);
}

int fl_sgetrs(const float* LU, int64_t n, int64_t lda, const void* aux, float* B, int64_t nrhs, int64_t ldb,
              void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return sgetrs(LU, n, lda, aux, B, nrhs, ldb, workspace, workspace_bytes, as_stream(stream));
}

int fl_dgetrs(const double* LU, int64_t n, int64_t lda, const void* aux, double* B, int64_t nrhs, int64_t ldb,
              void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return dgetrs(LU, n, lda, aux, B, nrhs, ldb, workspace, workspace_bytes, as_stream(stream));
}

size_t fl_refine_workspace_size(int64_t n, int64_t nrhs) { return refine_workspace_bytes(n, nrhs); }

int fl_refine(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32, int64_t n,
              const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
              double* resid_out, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    return refine_impl(A64, lda64, LU32, lda32, aux32, n, b, x, nrhs, ldb, tol, max_iter, iters_out, resid_out,
                       nullptr, workspace, workspace_bytes, as_stream(stream));
}

int fl_refine_frame(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32, int64_t n,
                    const double* frame, const double* b, double* x, int64_t nrhs, int64_t ldb, double tol,
                    int max_iter, int* iters_out, double* resid_out, void* workspace, size_t workspace_bytes,
                    void* stream)
{
    clear_error();
    if (!frame) return set_error(FL_EINVAL, "fl_refine_frame needs a frame");
    return refine_impl(A64, lda64, LU32, lda32, aux32, n, b, x, nrhs, ldb, tol, max_iter, iters_out, resid_out,
                       frame, workspace, workspace_bytes, as_stream(stream));
}

int fl_refine_steps(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32,
                    int64_t n, const double* b, double* x, int64_t nrhs, int64_t ldb, int steps, double* resid_dev,
                    void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (!resid_dev) return set_error(FL_EINVAL, "fl_refine_steps needs resid_dev");
    if (steps < 0) return set_error(FL_EINVAL, "fl_refine_steps: negative step count");
    return refine_impl(A64, lda64, LU32, lda32, aux32, n, b, x, nrhs, ldb, 0.0, steps, nullptr, nullptr, nullptr,
                       workspace, workspace_bytes, as_stream(stream), nullptr, 0, resid_dev);
}

int fl_pad_solve(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                 int64_t nrhs, void* stream)
{
    clear_error();
    return pad_solve_impl(K64, K32, n_pad, n_dev, rhs, b, nrhs, as_stream(stream));
}

// ---------------------------------------------------------------- grouped calls
namespace {
int check_group(int64_t count, int64_t slab_bytes)
{
    FL_CHECK(count >= 1 && count <= 65535, "group size %lld outside [1, 65535]", (long long)count);
    FL_CHECK(count == 1 || (slab_bytes > 0 && slab_bytes % 256 == 0),
             "slab_bytes must be a positive multiple of 256 (got %lld)", (long long)slab_bytes);
    return FL_OK;
}
}  // namespace

int fl_pad_solve_group(double* K64, float* K32, int64_t n_pad, const double* n_dev, const double* rhs, double* b,
                       int64_t nrhs, int64_t count, int64_t slab_bytes, void* stream)
{
    clear_error();
    if (int e = check_group(count, slab_bytes)) return e;
    BatchScope scope(count, slab_bytes);
    return pad_solve_impl(K64, K32, n_pad, n_dev, rhs, b, nrhs, as_stream(stream));
}

int fl_sgetrf_async_group(float* A, int64_t n, int64_t lda, int32_t* ipiv, void* aux, unsigned int* status,
                          int64_t count, int64_t slab_bytes, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (int e = check_group(count, slab_bytes)) return e;
    BatchScope scope(count, slab_bytes);
    return sgetrf_async(A, n, lda, ipiv, aux, status, FL_LU_FP32, workspace, workspace_bytes, as_stream(stream));
}

int fl_refine_steps_group(const double* A64, int64_t lda64, const float* LU32, int64_t lda32, const void* aux32,
                          int64_t n, const double* b, double* x, int64_t nrhs, int64_t ldb, int steps,
                          double* resid_dev, int64_t count, int64_t slab_bytes, void* workspace,
                          size_t workspace_bytes, void* stream)
{
    clear_error();
    if (!resid_dev) return set_error(FL_EINVAL, "fl_refine_steps_group needs resid_dev");
    if (steps < 0) return set_error(FL_EINVAL, "fl_refine_steps_group: negative step count");
    if (int e = check_group(count, slab_bytes)) return e;
    BatchScope scope(count, slab_bytes);
    return refine_impl(A64, lda64, LU32, lda32, aux32, n, b, x, nrhs, ldb, 0.0, steps, nullptr, nullptr, nullptr,
                       workspace, workspace_bytes, as_stream(stream), nullptr, 0, resid_dev);
}

int fl_sgetri(const float* LU, int64_t n, int64_t lda, const void* aux, float* Ainv, int64_t ldi, void* stream)
{
    clear_error();
    return getri_impl(LU, n, lda, aux, Ainv, ldi, as_stream(stream));
}

int fl_refine_inv(const double* A64, int64_t lda64, const float* Ainv32, int64_t ldi, int64_t n, const double* frame,
                  const double* b, double* x, int64_t nrhs, int64_t ldb, double tol, int max_iter, int* iters_out,
                  double* resid_out, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (!Ainv32 && n > 0) return set_error(FL_EINVAL, "fl_refine_inv needs the inverse");
    return refine_impl(A64, lda64, nullptr, 0, nullptr, n, b, x, nrhs, ldb, tol, max_iter, iters_out, resid_out,
                       frame, workspace, workspace_bytes, as_stream(stream), Ainv32, ldi);
}

size_t fl_gmres_workspace_size(int64_t N, int restart, int inner_precision)
{
    return gmres_workspace_bytes(N, restart, inner_precision, 0);
}

int fl_gmres(const fl_kernel* kernel, int64_t N, const double* X, const double* eps, int64_t eps_stride, double mu_re,
             const double* U, double* f, double rtol, int restart, int maxiter, int inner_precision, int* iters_out,
             double* resid_out, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    return gmres_impl(kernel->kind, kernel->alpha, N, X, eps, eps_stride, mu_re, U, f, rtol, restart, maxiter,
                      inner_precision, iters_out, resid_out, workspace, workspace_bytes, as_stream(stream), 0);
}

size_t fl_gmres_bj_workspace_size(int64_t N, int restart, int block_points)
{
    return gmres_workspace_bytes(N, restart, FL_FP64, block_points);
}

int fl_gmres_bj(const fl_kernel* kernel, int64_t N, const double* X, const double* eps, int64_t eps_stride,
                double mu_re, const double* U, double* f, double rtol, int restart, int maxiter, int block_points,
                int* iters_out, double* resid_out, void* workspace, size_t workspace_bytes, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    if (block_points <= 0) return set_error(FL_EINVAL, "fl_gmres_bj needs a positive block size");
    return gmres_impl(kernel->kind, kernel->alpha, N, X, eps, eps_stride, mu_re, U, f, rtol, restart, maxiter, FL_FP64,
                      iters_out, resid_out, workspace, workspace_bytes, as_stream(stream), block_points);
}

int fl_block_lu(int64_t m, int64_t nblk, double* blocks, int32_t* ipiv, int32_t* perm, int32_t* info, void* stream)
{
    clear_error();
    if (nblk > 0 && !(blocks && ipiv && perm && info)) return set_error(FL_EINVAL, "null pointer argument");
    return block_lu_launch(m, nblk, blocks, ipiv, perm, info, as_stream(stream));
}

int fl_block_apply(int64_t N, int block_points, const double* blocks, const int32_t* perm, const double* v, double* z,
                   void* stream)
{
    clear_error();
    if (N > 0 && !(blocks && perm && v && z)) return set_error(FL_EINVAL, "null pointer argument");
    return block_apply_launch(N, block_points, blocks, perm, v, z, as_stream(stream));
}

int fl_assemble_blocks(const fl_kernel* kernel, int64_t N, int block_points, const double* X, const double* eps,
                       int64_t eps_stride, double mu_re, double* blocks, void* stream)
{
    clear_error();
    if (int s = check_kernel(kernel)) return s;
    FL_CHECK(mu_re != 0.0, "mu must be nonzero");
    return assemble_blocks_launch(kernel->kind, kernel->alpha, N, block_points, X, eps, eps_stride,
                                  kernel_scale(kernel->kind) / mu_re, blocks, as_stream(stream));
}

int fl_sgemm(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
             float beta, float* C, int64_t ldc, void* stream)
{
    clear_error();
    FL_CHECK(lda >= K && ldb >= N && ldc >= N, "leading dimension too small");
    return gemm_launch<float>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, as_stream(stream));
}

int fl_dgemm(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, void* stream)
{
    clear_error();
    FL_CHECK(lda >= K && ldb >= N && ldc >= N, "leading dimension too small");
    return gemm_launch<double>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, as_stream(stream));
}

int fl_swim_rhs(int64_t N, const double* x, const double* x0_host, const double* v, double* R, void* stream)
{
    clear_error();
    return swim_rhs_launch(N, x, x0_host, v, R, as_stream(stream));
}

int fl_swim_schur(int64_t N, const double* Y, const double* count, const double* Xsum, double* f, double* UW,
                  double* S, void* stream)
{
    clear_error();
    return swim_schur_launch(N, Y, count, Xsum, f, UW, S, as_stream(stream));
}

int fl_nearest(int64_t Q, int64_t N, const double* quad, const double* force, int64_t* nearest, void* stream)
{
    clear_error();
    return nearest_launch(Q, N, quad, force, nearest, as_stream(stream));
}

}  // extern "C"
