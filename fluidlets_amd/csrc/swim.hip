// Free-swimmer bordered system by Schur complement on the device.
//
// The reference assembles and solves the (3N+6)^2 matrix of
// src/fluidlets/swimming.py:90-101 with numpy.linalg.solve:
//   [ K   C ] [f]   [v]        C = [-I | [x_n - x0]x]  (3N x 6),
//   [ B   0 ] [w] = [0],       B = [count_n I ; [Xsum_n]x]  (6 x 3N),  w = [U; Omega].
// Here K is factored once and Y = K^{-1} [v | C] (seven right-hand sides) is
// solved with the LU of the resistance path; these kernels form the rest:
//   swim_rhs:    the 3N x 7 right-hand side [v | C] (row-major, [point, component] rows)
//   swim_schur:  S = B Y_C (6 x 6) and g = B Y_v in FP64 with a fixed-order
//                reduction, then S w = g by Gaussian elimination with partial
//                pivoting (one lane, FP64)
//   swim_forces: f = Y_v - Y_C w
#include "fl_internal.h"

namespace fl {
namespace {

constexpr int WG = 256;

// One lane per point n writes rows 3n, 3n+1, 3n+2 of R = [v | -I | [x_n - x0]x]
// ([a]x b = a x b; its rows are fixed at compile time, so no lane branches).
__global__ __launch_bounds__(WG) void swim_rhs_kernel(int64_t N, const double* x, double x00, double x01, double x02,
                                                      const double* v, double* R)
{
    const int64_t n = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (n >= N) return;
    const double r0 = x[n] - x00, r1 = x[N + n] - x01, r2 = x[2 * N + n] - x02;
    const double cx[3][3] = {{0.0, -r2, r1}, {r2, 0.0, -r0}, {-r1, r0, 0.0}};
    double* row = R + 3 * n * 7;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        row[a * 7 + 0] = v[a * N + n];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            row[a * 7 + 1 + b] = a == b ? -1.0 : 0.0;
            row[a * 7 + 4 + b] = cx[a][b];
        }
    }
}

// One workgroup.  BY[i][c], i < 6 rows of B, c < 7 columns of Y.
__global__ __launch_bounds__(WG) void swim_schur_kernel(int64_t N, const double* Y, const double* count,
                                                        const double* Xsum, double* UW, double* S_out)
{
    __shared__ double part[WG][42];
    __shared__ double BY[42];
    const int tid = threadIdx.x;
    double acc[42];
#pragma unroll
    for (int k = 0; k < 42; ++k) acc[k] = 0.0;
    for (int64_t n = tid; n < N; n += WG) {
        const double cn = count[n];
        const double xs[3] = {Xsum[n], Xsum[N + n], Xsum[2 * N + n]};
        const double* y = Y + 3 * n * 7;
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            const double y0 = y[c], y1 = y[7 + c], y2 = y[14 + c];
            acc[0 * 7 + c] += cn * y0;
            acc[1 * 7 + c] += cn * y1;
            acc[2 * 7 + c] += cn * y2;
            // torque rows: Xsum x Y_n
            acc[3 * 7 + c] += xs[1] * y2 - xs[2] * y1;
            acc[4 * 7 + c] += xs[2] * y0 - xs[0] * y2;
            acc[5 * 7 + c] += xs[0] * y1 - xs[1] * y0;
        }
    }
#pragma unroll
    for (int k = 0; k < 42; ++k) part[tid][k] = acc[k];
    __syncthreads();
    if (tid < 42) {
        double s = 0.0;
        for (int t = 0; t < WG; ++t) s += part[t][tid];
        BY[tid] = s;
    }
    __syncthreads();
    if (tid == 0) {
        // S = BY[:, 1:7], g = BY[:, 0]; Gaussian elimination with partial pivoting
        double M[6][7];
        for (int i = 0; i < 6; ++i) {
            for (int j = 0; j < 6; ++j) M[i][j] = BY[i * 7 + 1 + j];
            M[i][6] = BY[i * 7];
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) S_out[i * 6 + j] = M[i][j];
        for (int c = 0; c < 6; ++c) {
            int p = c;
            for (int r = c + 1; r < 6; ++r)
                if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
            if (p != c)
                for (int j = 0; j < 7; ++j) {
                    double t = M[c][j];
                    M[c][j] = M[p][j];
                    M[p][j] = t;
                }
            const double piv = M[c][c];
            for (int r = c + 1; r < 6; ++r) {
                const double l = piv != 0.0 ? M[r][c] / piv : 0.0;
                for (int j = c; j < 7; ++j) M[r][j] -= l * M[c][j];
            }
        }
        double w[6];
        for (int i = 5; i >= 0; --i) {
            double s = M[i][6];
            for (int j = i + 1; j < 6; ++j) s -= M[i][j] * w[j];
            w[i] = M[i][i] != 0.0 ? s / M[i][i] : 0.0;
        }
        for (int i = 0; i < 6; ++i) UW[i] = w[i];
    }
}

__global__ __launch_bounds__(WG) void swim_forces_kernel(int64_t N, const double* Y, const double* UW, double* f)
{
    const int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (i >= 3 * N) return;
    const double* y = Y + i * 7;
    double s = y[0];
#pragma unroll
    for (int c = 0; c < 6; ++c) s -= y[1 + c] * UW[c];
    f[i] = s;
}

}  // namespace

int swim_rhs_launch(int64_t N, const double* x, const double* x0, const double* v, double* R, hipStream_t st)
{
    FL_CHECK(N >= 0, "negative N");
    FL_CHECK(x0 != nullptr, "x0 is NULL");
    if (N == 0) return FL_OK;
    FL_CHECK(x && v && R, "null pointer argument");
    hipLaunchKernelGGL(swim_rhs_kernel, dim3((unsigned)((N + WG - 1) / WG)), dim3(WG), 0, st, N, x, x0[0], x0[1],
                       x0[2], v, R);
    FL_LAUNCHED();
    return FL_OK;
}

int swim_schur_launch(int64_t N, const double* Y, const double* count, const double* Xsum, double* f, double* UW,
                      double* S, hipStream_t st)
{
    FL_CHECK(N > 0, "N must be positive");
    FL_CHECK(Y && count && Xsum && f && UW && S, "null pointer argument");
    hipLaunchKernelGGL(swim_schur_kernel, dim3(1), dim3(WG), 0, st, N, Y, count, Xsum, UW, S);
    FL_LAUNCHED();
    hipLaunchKernelGGL(swim_forces_kernel, dim3((unsigned)((3 * N + WG - 1) / WG)), dim3(WG), 0, st, N, Y,
                       (const double*)UW, f);
    FL_LAUNCHED();
    return FL_OK;
}

}  // namespace fl
