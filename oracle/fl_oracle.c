/*
 * fl_oracle.c -- CPU restatement of the Fluidlets kernel sums and dense fills.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and
 * only as the checker or the timed CPU baseline.  The product (fluidlets_amd/)
 * never links, loads or calls it.
 *
 * It follows the reference arithmetic form by form (same per-target loop,
 * same sequential source order, the same pow()-based expressions, no
 * fast-math, no FMA contraction: build with -O3 -ffp-contract=off):
 *
 *   stokeslet_block   src/fluidlets/kernels/stokeslet3d.py:22-50
 *   image_block       src/fluidlets/kernels/images.py:26-78
 *   wall_block        src/fluidlets/kernels/images.py:81-89
 *   brink_f1/f2       src/fluidlets/kernels/brinkmanlet3d.py:43-71
 *   brinkman_block    src/fluidlets/kernels/brinkmanlet3d.py:74-104
 *   flo_apply         src/fluidlets/core/matvec.py:10-21   (_apply)
 *   flo_fill          src/fluidlets/core/assemble.py:15-24 (_fill)
 *   flo_fill_nearest  src/fluidlets/core/assemble.py:49-58 (_fill_nearest)
 *
 * Points are AoS (N,3) row-major, as in the reference ([point, component]).
 * Complex strengths are passed as separate real and imaginary arrays: the
 * block is real, so the reference's complex accumulation is two independent
 * real accumulations in the same order.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FLO_STOKESLET 0
#define FLO_WALL 1
#define FLO_BRINKMAN 2

static const double PI = 3.141592653589793;

/* kernels/stokeslet3d.py:22-29 */
static double H1s(double r2, double eps2) { return (r2 + 2.0 * eps2) / (8.0 * PI * pow(r2 + eps2, 1.5)); }
static double H2s(double r2, double eps2) { return 1.0 / (8.0 * PI * pow(r2 + eps2, 1.5)); }

/* kernels/stokeslet3d.py:32-50 */
static void stokeslet_block(const double* x, const double* x0, double eps, double* out)
{
    double d0 = x[0] - x0[0], d1 = x[1] - x0[1], d2 = x[2] - x0[2];
    double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    double eps2 = eps * eps;
    double h1 = H1s(r2, eps2), h2 = H2s(r2, eps2);
    out[0] = h1 + h2 * d0 * d0;
    out[1] = h2 * d0 * d1;
    out[2] = h2 * d0 * d2;
    out[3] = out[1];
    out[4] = h1 + h2 * d1 * d1;
    out[5] = h2 * d1 * d2;
    out[6] = out[2];
    out[7] = out[5];
    out[8] = h1 + h2 * d2 * d2;
}

/* kernels/images.py:26-78: image part of the wall system, mu = 1 */
static void image_block(const double* x, const double* x0, double eps, double* out)
{
    double h0 = x0[2];
    double d0 = x[0] - x0[0], d1 = x[1] - x0[1], d2 = x[2] + h0;
    double eps2 = eps * eps;
    double Q = d0 * d0 + d1 * d1 + d2 * d2 + eps2;
    double q1 = pow(Q, 0.5), q3 = pow(Q, 1.5), q5 = pow(Q, 2.5);
    double H1 = 0.5 / q1 + 0.5 * eps2 / q3;
    double H2 = 0.5 / q3;
    double Hdp1 = -h0 * h0 * (1.0 / q3 - 3.0 * eps2 / q5);
    double Hdp2 = -h0 * h0 * (-3.0 / q5);
    double Hdb1 = 2.0 * h0 * (-1.5 * eps2 / q5 - 0.5 / q3);
    double Hdb2 = 2.0 * h0 * (0.5 / q3);
    double Hdb3 = 2.0 * h0 * (-1.5 / q5);
    double Hr1 = h0 * 3.0 * eps2 / q5;
    double d[3] = {d0, d1, d2};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double delta = (i == j) ? 1.0 : 0.0;
            double a = -(delta * H1 + d[i] * d[j] * H2);
            double sgn = (j == 2) ? 1.0 : -1.0;
            a += sgn * (delta * Hdp1 + d[i] * d[j] * Hdp2);
            out[3 * i + j] = a;
        }
    out[0] -= d2 * (Hdb2 + d0 * d0 * Hdb3);
    out[1] -= d0 * d1 * d2 * Hdb3;
    out[2] += d0 * (Hdb2 + d2 * d2 * Hdb3);
    out[3] -= d1 * d0 * d2 * Hdb3;
    out[4] -= d2 * (Hdb2 + d1 * d1 * Hdb3);
    out[5] += d1 * (Hdb2 + d2 * d2 * Hdb3);
    out[6] -= d0 * (Hdb1 + d2 * d2 * Hdb3);
    out[7] -= d1 * (Hdb1 + d2 * d2 * Hdb3);
    out[8] += d2 * (Hdb1 + 2.0 * Hdb2 + d2 * d2 * Hdb3);
    out[0] += d2 * Hr1;
    out[4] += d2 * Hr1;
    out[6] -= d0 * Hr1;
    out[7] -= d1 * Hr1;
    for (int k = 0; k < 9; ++k) out[k] /= 4.0 * PI;
}

/* kernels/images.py:81-89 */
static void wall_block(const double* x, const double* x0, double eps, double* out)
{
    double tmp[9];
    image_block(x, x0, eps, out);
    stokeslet_block(x, x0, eps, tmp);
    for (int k = 0; k < 9; ++k) out[k] += tmp[k];
}

/* kernels/brinkmanlet3d.py:43-49: 24-term Taylor coefficients below z = 0.5 */
#define FLO_NTERMS 24
static double C1[FLO_NTERMS], C2[FLO_NTERMS];
static int coeffs_ready = 0;

static double inv_factorial(int n)
{
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= (double)k;
    return 1.0 / f;
}

void flo_init(void)
{
    if (coeffs_ready) return;
    for (int n = 0; n < FLO_NTERMS; ++n) {
        double s = (n % 2 == 0) ? 1.0 : -1.0;
        C1[n] = s * (inv_factorial(n + 2) - inv_factorial(n + 1) + inv_factorial(n));
        C2[n] = -s * (3.0 * inv_factorial(n + 2) - 3.0 * inv_factorial(n + 1) + inv_factorial(n));
    }
    coeffs_ready = 1;
}

/* kernels/brinkmanlet3d.py:52-60 */
double flo_brink_f1(double z)
{
    if (z < 0.5) {
        double s = 0.0;
        for (int n = FLO_NTERMS - 1; n >= 0; --n) s = s * z + C1[n];
        return s;
    }
    return exp(-z) * (1.0 + 1.0 / z + 1.0 / (z * z)) - 1.0 / (z * z);
}

/* kernels/brinkmanlet3d.py:63-71 */
double flo_brink_f2(double z)
{
    if (z < 0.5) {
        double s = 0.0;
        for (int n = FLO_NTERMS - 1; n >= 0; --n) s = s * z + C2[n];
        return s;
    }
    return 3.0 / (z * z) - exp(-z) * (1.0 + 3.0 / z + 3.0 / (z * z));
}

/* kernels/brinkmanlet3d.py:74-104 */
static void brinkman_block(const double* x, const double* x0, double eps, double alpha, double* out)
{
    double d0 = x[0] - x0[0], d1 = x[1] - x0[1], d2 = x[2] - x0[2];
    double r2 = d0 * d0 + d1 * d1 + d2 * d2, eps2 = eps * eps;
    double R = sqrt(r2 + eps2);
    double z = alpha * R;
    double h1 = flo_brink_f1(z) / (4.0 * PI * R);
    double h2 = flo_brink_f2(z) / (4.0 * PI * R * R * R);
    double H1 = h1 + eps2 * h2, H2 = h2;
    out[0] = H1 + H2 * d0 * d0;
    out[1] = H2 * d0 * d1;
    out[2] = H2 * d0 * d2;
    out[3] = out[1];
    out[4] = H1 + H2 * d1 * d1;
    out[5] = H2 * d1 * d2;
    out[6] = out[2];
    out[7] = out[5];
    out[8] = H1 + H2 * d2 * d2;
}

static void block(int kind, double alpha, const double* x, const double* x0, double eps, double* out)
{
    if (kind == FLO_STOKESLET) stokeslet_block(x, x0, eps, out);
    else if (kind == FLO_WALL) wall_block(x, x0, eps, out);
    else brinkman_block(x, x0, eps, alpha, out);
}

/* Single 3x3 block, for tests of the block functions themselves. */
void flo_block(int kind, double alpha, const double* x, const double* x0, double eps, double* out)
{
    flo_init();
    block(kind, alpha, x, x0, eps, out);
}

void flo_image_block(const double* x, const double* x0, double eps, double* out)
{
    image_block(x, x0, eps, out);
}

void flo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int flo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* core/matvec.py:10-21.  out_re/out_im are (M,3), zero on entry. */
void flo_apply(int kind, double alpha, int64_t M, int64_t N,
               const double* targets, const double* sources, const double* eps,
               const double* f_re, const double* f_im, double* out_re, double* out_im)
{
    flo_init();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        double G[9];
        for (int64_t j = 0; j < N; ++j) {
            block(kind, alpha, targets + 3 * i, sources + 3 * j, eps[j], G);
            for (int a = 0; a < 3; ++a) {
                double acc = 0.0;
                for (int b = 0; b < 3; ++b) acc += G[3 * a + b] * f_re[3 * j + b];
                out_re[3 * i + a] += acc;
                if (f_im) {
                    double acci = 0.0;
                    for (int b = 0; b < 3; ++b) acci += G[3 * a + b] * f_im[3 * j + b];
                    out_im[3 * i + a] += acci;
                }
            }
        }
    }
}

/* core/assemble.py:15-24.  K is (3M, 3N) row-major. */
void flo_fill(int kind, double alpha, int64_t M, int64_t N,
              const double* targets, const double* sources, const double* eps, double* K)
{
    flo_init();
    const int64_t ld = 3 * N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        double G[9];
        for (int64_t j = 0; j < N; ++j) {
            block(kind, alpha, targets + 3 * i, sources + 3 * j, eps[j], G);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) K[(3 * i + a) * ld + 3 * j + b] = G[3 * a + b];
        }
    }
}

/* core/assemble.py:49-58.  K is (3M, 3*n_force) row-major and zero on entry. */
void flo_fill_nearest(int kind, double alpha, int64_t M, int64_t Q, int64_t n_force,
                      const double* targets, const double* quad, const int64_t* nearest,
                      const double* eps, double* K)
{
    flo_init();
    const int64_t ld = 3 * n_force;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        double G[9];
        for (int64_t k = 0; k < Q; ++k) {
            block(kind, alpha, targets + 3 * i, quad + 3 * k, eps[k], G);
            int64_t j = nearest[k];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) K[(3 * i + a) * ld + 3 * j + b] += G[3 * a + b];
        }
    }
}
