// Restarted GMRES on the device with the matrix-free kernel sum as operator.
//
// Replaces scipy.sparse.linalg.gmres over the evaluate LinearOperator at
// src/fluidlets/core/solve.py:68-79.  Vectors are the (3, N) structure-of-
// arrays layout of the kernel sums (a fixed permutation of the reference's
// [point, component] order, which GMRES is invariant to).  Arnoldi uses
// classical Gram-Schmidt applied twice (two multi-dot/multi-axpy passes per
// step); dot products are reduced in a fixed order, so runs are
// deterministic.  The (m+1) x m Hessenberg least-squares problem is solved
// on the host with Givens rotations, and the Givens residual estimate stops
// a cycle early, as scipy does; every restart recomputes the true residual.
//
// Flexible variant (inner_precision = FL_FP32C): FGMRES(m) in FP64 whose
// preconditioner z = M^{-1} v is one inner GMRES(m_in) cycle with FP32C
// matvecs; the outer matvecs, residuals and solution stay FP64.
//
// Block-Jacobi variant (block_points = P > 0): the same right-preconditioned
// loop with z = blockdiag(K_b)^{-1} v, K_b the FP64 matrix of body b (points
// b P .. b P + P - 1), assembled and factored once (bjacobi.hip).
#include "fl_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fl {
namespace {

constexpr int WG = 256;
constexpr int kInnerRestart = 20;  // inner FP32C cycle length of the flexible variant

// out[i] = sum_k V[i][k] w[k] for i < m (one workgroup per i, fixed order).
__global__ __launch_bounds__(WG) void multidot_kernel(const double* V, int64_t ldv, const double* w, int64_t n,
                                                      double* out)
{
    __shared__ double s[WG];
    const double* v = V + (int64_t)blockIdx.x * ldv;
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += WG) acc += v[k] * w[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = WG / 2; h > 0; h /= 2) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// w[k] += alpha * sum_{i<m} V[i][k] c[i]
__global__ __launch_bounds__(WG) void multiaxpy_kernel(const double* V, int64_t ldv, const double* c, int m,
                                                       double alpha, double* w, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (k >= n) return;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += V[i * ldv + k] * c[i];
    w[k] += alpha * s;
}

// y = a x + b y  (y is not read when b == 0, so it may start uninitialised)
__global__ __launch_bounds__(WG) void axpby_kernel(double a, const double* x, double b, double* y, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (k < n) y[k] = b == 0.0 ? a * x[k] : a * x[k] + b * y[k];
}

int64_t blocks(int64_t n) { return (n + WG - 1) / WG; }

struct Op {
    int kind;
    double alpha;
    int64_t N;
    const double* X;
    const double* eps;
    int64_t eps_stride;
    double w_re;
    void* ews;
    size_t ews_bytes;
    hipStream_t st;
    int apply(const double* v, double* out, int precision) const
    {
        return eval_launch(kind, alpha, N, N, X, X, eps, eps_stride, v, nullptr, w_re, 0.0, out, nullptr, precision,
                           ews, ews_bytes, st);
    }
};

struct Scratch {
    double* dots;  // device, >= 2 * (m + 2)
    std::vector<double> host;
};

int dot_host(const double* a, const double* b, int64_t n, Scratch& s, hipStream_t st, double& out)
{
    hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(WG), 0, st, a, (int64_t)0, b, n, s.dots);
    FL_LAUNCHED();
    FL_HIP(hipMemcpyAsync(&out, s.dots, sizeof(double), hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    return FL_OK;
}

// One Arnoldi step: orthogonalise w against V[0..j] twice (CGS2); returns
// h[0..j] and ||w|| on the host; w is left unnormalised.
int arnoldi(const double* V, int64_t ldv, int j, double* w, int64_t n, Scratch& s, hipStream_t st,
            std::vector<double>& h, double& wnorm)
{
    h.assign(j + 1, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(multidot_kernel, dim3(j + 1), dim3(WG), 0, st, V, ldv, (const double*)w, n, s.dots);
        FL_LAUNCHED();
        hipLaunchKernelGGL(multiaxpy_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, V, ldv, (const double*)s.dots,
                           j + 1, -1.0, w, n);
        FL_LAUNCHED();
        FL_HIP(hipMemcpyAsync(s.host.data(), s.dots, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        for (int i = 0; i <= j; ++i) h[i] += s.host[i];
    }
    if (int e = dot_host(w, w, n, s, st, wnorm)) return e;
    wnorm = std::sqrt(wnorm);
    return FL_OK;
}

// Givens least-squares state of one cycle.
struct LSQ {
    int m;
    std::vector<double> H;  // (m+1) x m column-major: H[c * (m+1) + r]
    std::vector<double> cs, sn, g;
    explicit LSQ(int m_) : m(m_), H((m_ + 1) * m_, 0.0), cs(m_), sn(m_), g(m_ + 1, 0.0) {}
    void reset(double beta)
    {
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
    }
    // adds column j (h[0..j], h_{j+1,j} = hn) and returns |residual estimate|
    double add(int j, const std::vector<double>& h, double hn)
    {
        double* col = &H[(size_t)j * (m + 1)];
        for (int i = 0; i <= j; ++i) col[i] = h[i];
        col[j + 1] = hn;
        for (int i = 0; i < j; ++i) {
            double a = cs[i] * col[i] + sn[i] * col[i + 1];
            double b = -sn[i] * col[i] + cs[i] * col[i + 1];
            col[i] = a;
            col[i + 1] = b;
        }
        double den = std::hypot(col[j], col[j + 1]);
        cs[j] = den > 0 ? col[j] / den : 1.0;
        sn[j] = den > 0 ? col[j + 1] / den : 0.0;
        col[j] = den;
        col[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        return std::fabs(g[j + 1]);
    }
    // y solving the k x k upper triangular system R y = g
    std::vector<double> solve(int k) const
    {
        std::vector<double> y(k);
        for (int i = k - 1; i >= 0; --i) {
            double s = g[i];
            for (int c = i + 1; c < k; ++c) s -= H[(size_t)c * (m + 1) + i] * y[c];
            y[i] = H[(size_t)i * (m + 1) + i] != 0 ? s / H[(size_t)i * (m + 1) + i] : 0.0;
        }
        return y;
    }
};

// One inner GMRES(m) cycle from zero on rhs v (precision prec): z ~= A^{-1} v.
int inner_cycle(const Op& op, const double* v, double* z, int m, double* V, double* w, Scratch& s, int prec,
                int64_t n, int& its)
{
    hipStream_t st = op.st;
    double beta = 0.0;
    if (int e = dot_host(v, v, n, s, st, beta)) return e;
    beta = std::sqrt(beta);
    FL_HIP(hipMemsetAsync(z, 0, n * sizeof(double), st));
    if (beta == 0.0) return FL_OK;
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0 / beta, v, 0.0, V, n);
    FL_LAUNCHED();
    LSQ L(m);
    L.reset(beta);
    std::vector<double> h;
    int k = 0;
    for (int j = 0; j < m; ++j) {
        if (int e = op.apply(V + (int64_t)j * n, w, prec)) return e;
        double hn = 0.0;
        if (int e = arnoldi(V, n, j, w, n, s, st, h, hn)) return e;
        double est = L.add(j, h, hn);
        ++its;
        k = j + 1;
        if (hn == 0.0 || est <= 1e-3 * beta) break;
        if (j + 1 < m) {
            hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0 / hn, (const double*)w, 0.0,
                               V + (int64_t)(j + 1) * n, n);
            FL_LAUNCHED();
        }
    }
    std::vector<double> y = L.solve(k);
    FL_HIP(hipMemcpyAsync(s.dots, y.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(multiaxpy_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, (const double*)V, n,
                       (const double*)s.dots, k, 1.0, z, n);
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace

// Block-Jacobi storage: the factored blocks, then ipiv, perm and info.
size_t bj_bytes(int64_t N, int P)
{
    if (P <= 0 || N % P != 0) return 0;
    const int64_t m = 3 * (int64_t)P, nblk = N / P;
    return align_up((size_t)nblk * m * m * sizeof(double), 256) + 2 * align_up((size_t)nblk * m * sizeof(int32_t), 256) +
           align_up((size_t)nblk * sizeof(int32_t), 256);
}

size_t gmres_workspace_bytes(int64_t N, int restart, int inner_precision, int block_points)
{
    const int64_t n = 3 * N;
    const size_t vec = align_up((size_t)n * sizeof(double), 256);
    size_t b = (size_t)(restart + 1) * vec + 2 * vec + 4096;  // V, w, r, dots
    if (inner_precision == FL_FP32C) b += (size_t)restart * vec + (size_t)(kInnerRestart + 1) * vec + vec;
    else if (block_points > 0) b += (size_t)restart * vec + bj_bytes(N, block_points);
    b += align_up(std::max({eval_workspace_size(N, N, 1, inner_precision), eval_workspace_size(N, N, 1, FL_FP64),
                            eval_sym_workspace_size(N, 1, FL_FP64)}),
                  256);
    return b;
}

int gmres_impl(int kind, double alpha, int64_t N, const double* X, const double* eps, int64_t eps_stride,
               double mu_re, const double* U, double* f, double rtol, int restart, int maxiter, int inner_precision,
               int* iters_out, double* resid_out, void* ws, size_t ws_bytes, hipStream_t st, int block_points)
{
    FL_CHECK(N >= 0 && restart >= 1 && maxiter >= 1, "invalid GMRES sizes");
    FL_CHECK(inner_precision == FL_FP64 || inner_precision == FL_FP32C, "unknown inner precision");
    FL_CHECK(mu_re != 0.0, "mu must be nonzero");
    FL_CHECK(block_points >= 0, "negative block size");
    FL_CHECK(block_points == 0 || inner_precision == FL_FP64, "block-Jacobi runs with FP64 matvecs");
    FL_CHECK(block_points == 0 || (N % block_points == 0 && 3 * (int64_t)block_points <= 512),
             "block-Jacobi needs N a multiple of the block size and at most 170 points per block");
    if (N == 0) return FL_OK;
    const size_t need = gmres_workspace_bytes(N, restart, inner_precision, block_points);
    if (ws_bytes < need || !ws) return set_error(FL_EWORKSPACE, "gmres needs %zu workspace bytes", need);
    const int64_t n = 3 * N;
    const size_t vec = align_up((size_t)n * sizeof(double), 256);
    char* p = static_cast<char*>(ws);
    double* V = reinterpret_cast<double*>(p);
    p += (size_t)(restart + 1) * vec;
    double* w = reinterpret_cast<double*>(p);
    p += vec;
    double* r = reinterpret_cast<double*>(p);
    p += vec;
    Scratch s;
    s.dots = reinterpret_cast<double*>(p);
    s.host.resize(512);
    p += 4096;
    FL_CHECK(restart + 2 <= 512, "restart too large");
    const bool flexible = inner_precision == FL_FP32C;
    const bool bj = block_points > 0;
    double *Z = nullptr, *Vin = nullptr, *win = nullptr, *KB = nullptr;
    int32_t *bj_ipiv = nullptr, *bj_perm = nullptr, *bj_info = nullptr;
    if (flexible) {
        Z = reinterpret_cast<double*>(p);
        p += (size_t)restart * vec;
        Vin = reinterpret_cast<double*>(p);
        p += (size_t)(kInnerRestart + 1) * vec;
        win = reinterpret_cast<double*>(p);
        p += vec;
    } else if (bj) {
        const int64_t m = 3 * (int64_t)block_points, nblk = N / block_points;
        Z = reinterpret_cast<double*>(p);
        p += (size_t)restart * vec;
        KB = reinterpret_cast<double*>(p);
        p += align_up((size_t)nblk * m * m * sizeof(double), 256);
        bj_ipiv = reinterpret_cast<int32_t*>(p);
        p += align_up((size_t)nblk * m * sizeof(int32_t), 256);
        bj_perm = reinterpret_cast<int32_t*>(p);
        p += align_up((size_t)nblk * m * sizeof(int32_t), 256);
        bj_info = reinterpret_cast<int32_t*>(p);
        p += align_up((size_t)nblk * sizeof(int32_t), 256);
    }
    Op op{kind, alpha, N, X, eps, eps_stride, kernel_scale(kind) / mu_re, nullptr, 0, st};
    op.ews = p;
    op.ews_bytes = std::max({eval_workspace_size(N, N, 1, FL_FP64), eval_workspace_size(N, N, 1, inner_precision),
                             eval_sym_workspace_size(N, 1, FL_FP64)});
    if (bj) {
        // the bodies' diagonal blocks of the operator (same scale), factored once
        const int64_t m = 3 * (int64_t)block_points, nblk = N / block_points;
        if (int e = assemble_blocks_launch(kind, alpha, N, block_points, X, eps, eps_stride, op.w_re, KB, st)) return e;
        if (int e = block_lu_launch(m, nblk, KB, bj_ipiv, bj_perm, bj_info, st)) return e;
        std::vector<int32_t> hinfo(nblk);
        FL_HIP(hipMemcpyAsync(hinfo.data(), bj_info, nblk * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FL_HIP(hipStreamSynchronize(st));
        for (int64_t q = 0; q < nblk; ++q)
            if (hinfo[q] != 0)
                return set_error(FL_ESINGULAR, "block-Jacobi: block %lld is singular (zero pivot %d)", (long long)q,
                                 hinfo[q]);
    }

    double bnorm = 0.0;
    if (int e = dot_host(U, U, n, s, st, bnorm)) return e;
    bnorm = std::sqrt(bnorm);
    if (bnorm == 0.0) {
        FL_HIP(hipMemsetAsync(f, 0, n * sizeof(double), st));
        if (iters_out) *iters_out = 0;
        if (resid_out) *resid_out = 0.0;
        return FL_OK;
    }
    const double target = rtol * bnorm;
    int its = 0;
    double rnorm = 0.0;
    std::vector<double> h;
    LSQ L(restart);
    for (int cycle = 0; cycle < maxiter; ++cycle) {
        // r = U - A f
        if (int e = op.apply(f, r, FL_FP64)) return e;
        hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0, U, -1.0, r, n);
        FL_LAUNCHED();
        if (int e = dot_host(r, r, n, s, st, rnorm)) return e;
        rnorm = std::sqrt(rnorm);
        if (rnorm <= target) break;
        hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0 / rnorm, (const double*)r, 0.0,
                           V, n);
        FL_LAUNCHED();
        L.reset(rnorm);
        int k = 0;
        for (int j = 0; j < restart; ++j) {
            const double* vj = V + (int64_t)j * n;
            if (flexible) {
                int inner_its = 0;
                if (int e = inner_cycle(op, vj, Z + (int64_t)j * n, kInnerRestart, Vin, win, s, FL_FP32C, n,
                                        inner_its))
                    return e;
                if (int e = op.apply(Z + (int64_t)j * n, w, FL_FP64)) return e;
            } else if (bj) {
                if (int e = block_apply_launch(N, block_points, KB, bj_perm, vj, Z + (int64_t)j * n, st)) return e;
                if (int e = op.apply(Z + (int64_t)j * n, w, FL_FP64)) return e;
            } else {
                if (int e = op.apply(vj, w, FL_FP64)) return e;
            }
            double hn = 0.0;
            if (int e = arnoldi(V, n, j, w, n, s, st, h, hn)) return e;
            const double est = L.add(j, h, hn);
            ++its;
            k = j + 1;
            if (hn == 0.0 || est <= target) break;
            if (j + 1 < restart) {
                hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0 / hn,
                                   (const double*)w, 0.0, V + (int64_t)(j + 1) * n, n);
                FL_LAUNCHED();
            }
        }
        std::vector<double> y = L.solve(k);
        FL_HIP(hipMemcpyAsync(s.dots, y.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(multiaxpy_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st,
                           (const double*)(flexible || bj ? Z : V),
                           n, (const double*)s.dots, k, 1.0, f, n);
        FL_LAUNCHED();
        rnorm = -1.0;
    }
    if (rnorm < 0.0) {
        if (int e = op.apply(f, r, FL_FP64)) return e;
        hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)blocks(n)), dim3(WG), 0, st, 1.0, U, -1.0, r, n);
        FL_LAUNCHED();
        if (int e = dot_host(r, r, n, s, st, rnorm)) return e;
        rnorm = std::sqrt(rnorm);
    }
    FL_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = its;
    if (resid_out) *resid_out = rnorm / bnorm;
    if (!(rnorm <= target))
        return set_error(FL_ENOCONV, "GMRES did not converge: relative residual %.3e after %d iterations",
                         rnorm / bnorm, its);
    return FL_OK;
}

}  // namespace fl
