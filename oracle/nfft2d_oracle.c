/*
 * nfft2d_oracle.c -- CPU restatement of the reference's 2D NFFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.  The
 * product (miscellaneous_amd/libnfft2d.so) never links or calls it.
 *
 * What it restates, step by step (paths relative to the reference root):
 *   orc_alpha            NFFT/NFFT2_2D.cu:27, BP_FBP_FFBP/NFFT1_2D.m:11
 *   orc_bessel_i0        the I0 behind Kernel_Bessel (NFFT/NFFT2_2D.cu:67) and
 *                        besseli(0,.) (BP_FBP_FFBP/NFFT1_2D.m:21-22); the
 *                        reference's own bessi0 lives in a header that is not in
 *                        the repository, so this is the textbook power series
 *                        sum_j (x^2/4)^j / (j!)^2.
 *   orc_window           spatial window phi_hat, NFFT/NFFT2_2D.cu:112-122 and
 *                        BP_FBP_FFBP/NFFT1_2D.m:34-41 (three branches).
 *   orc_window_deriv     d phi_hat / d d -- new (the reference never
 *                        differentiates); closed form plus a series near |d|=K.
 *   orc_nfft1_2d_{d,f}   type-2 (uniform -> non-uniform), BP_FBP_FFBP/NFFT1_2D.m
 *                        whole file: windows :17-42, scale+pad :48-49,
 *                        U = fft2(ifftshift(u)) :55, 13x13 gather :61-67.
 *   orc_nfft2_2d_{d,f}   type-1 (non-uniform -> uniform), NFFT/NFFT2_2D.cu
 *                        :329-411 (float) / :413-495 (double): Bessel table
 *                        :60-69, spread :163-218 with series_terms :105-130,
 *                        fftshift -> forward FFT -> fftshift :383-393,
 *                        Decimation_and_Scaling :74-100.  Also the MATLAB twin
 *                        NFFT/NFFT2/Matlab/NFFT2_2D.m.
 *   orc_ndft2_2d         direct NED sum of NFFT/NDFT2_2D.cu:28-29,67
 *                        (Kernel_Matrix_Filling + cublasZgemv OP_T).
 *   orc_ndft_type2(_dx)  direct type-2 sum y_l = sum_k f_k e^{-2 pi i k.x_l/N}
 *                        (sign pinned by NFFT/NFFT1/1D/Matlab/NFFT1_1D_BLAS.m:5-16)
 *                        and its analytic derivative in x.
 *
 * Conventions: complex arrays are interleaved (re, im).  Coordinates are in
 * grid-index units (period N), modes k in [-N/2, N/2) stored at index k+N/2,
 * 2-D arrays row-major [k1][k2].
 *
 * round_mode: 0 = rint (half to even, the CUDA path NFFT/NFFT2_2D.cu:168),
 *             1 = MATLAB round (half away from zero, NFFT1_2D.m:27).
 * mutation:   0 in every normal use.  Non-zero values plant a known mistake
 *             so that the tests can show their checks catch it:
 *             1 = flip the sign of the FFT exponent,
 *             2 = skip the ifftshift before the FFT,
 *             3 = window tap offset shifted by one grid cell.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643383279502884

double orc_alpha(int c) { return (2.0 - 1.0 / (double)c) * ORC_PI - 0.01; }

double orc_bessel_i0(double x)
{
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int j = 1; j < 1000; ++j) {
        term *= q / ((double)j * (double)j);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

/* f(P) = sinh(alpha sqrt(P)) / sqrt(P), continued through P <= 0 (sin branch,
 * value alpha at P == 0).  Power series sum_j alpha^(2j+1) P^j / (2j+1)! for
 * |P| < 1, closed forms otherwise. */
static double knab_f(double P, double alpha)
{
    if (fabs(P) < 1.0) {
        double a2 = alpha * alpha, term = alpha, sum = alpha;
        for (int j = 1; j < 60; ++j) {
            term *= a2 * P / ((2.0 * j) * (2.0 * j + 1.0));
            sum += term;
            if (fabs(term) < 1e-18 * fabs(sum)) break;
        }
        return sum;
    }
    if (P > 0) { double s = sqrt(P); return sinh(alpha * s) / s; }
    double s = sqrt(-P);
    return sin(alpha * s) / s;
}

/* f'(P) */
static double knab_fprime(double P, double alpha)
{
    if (fabs(P) < 1.0) {
        /* sum_{j>=1} j alpha^(2j+1) P^(j-1) / (2j+1)! */
        double a2 = alpha * alpha;
        double c = alpha * a2 / 6.0;           /* alpha^3/3!, j = 1 */
        double Pp = 1.0, sum = c;
        for (int j = 2; j < 60; ++j) {
            c *= a2 / ((2.0 * j) * (2.0 * j + 1.0));
            Pp *= P;
            double t = (double)j * c * Pp;
            sum += t;
            if (fabs(t) < 1e-18 * fabs(sum)) break;
        }
        return sum;
    }
    if (P > 0) {
        double s = sqrt(P);
        return (alpha * s * cosh(alpha * s) - sinh(alpha * s)) / (2.0 * s * s * s);
    }
    double s = sqrt(-P);
    /* f = sin(as)/s with s = sqrt(-P): df/dP = -(as cos(as) - sin(as)) / (2 s^3) */
    return -(alpha * s * cos(alpha * s) - sin(alpha * s)) / (2.0 * s * s * s);
}

/* Spatial window phi_hat(d) = (1/pi) sinh(alpha sqrt(K^2-d^2)) / sqrt(K^2-d^2)
 * (NFFT/NFFT2_2D.cu:112-116; BP_FBP_FFBP/NFFT1_2D.m:34-36).  The reference
 * evaluates the branches directly; near P = 0 the series above agrees with
 * them to rounding. */
double orc_window(double d, int K, double alpha)
{
    double P = (double)K * (double)K - d * d;
    if (P == 0.0) return alpha / ORC_PI;
    if (P > 0.0) { double s = sqrt(P); return sinh(alpha * s) / (s * ORC_PI); }
    double s = sqrt(-P);
    return sin(alpha * s) / (s * ORC_PI);
}

double orc_window_series(double d, int K, double alpha)
{
    double P = (double)K * (double)K - d * d;
    return knab_f(P, alpha) / ORC_PI;
}

double orc_window_deriv(double d, int K, double alpha)
{
    double P = (double)K * (double)K - d * d;
    return -2.0 * d * knab_fprime(P, alpha) / ORC_PI;
}

/* Spectral window phi(xi) = I0(K sqrt(alpha^2 - xi^2)) at xi = 2 pi k / (c N),
 * k = -N/2 .. N/2-1 (BP_FBP_FFBP/NFFT1_2D.m:17-22; NFFT/NFFT2_2D.cu:66-67). */
void orc_spectral_window(double *phi, int N, int K, int c)
{
    double alpha = orc_alpha(c);
    for (int i = 0; i < N; ++i) {
        double xi = 2.0 * ORC_PI * (double)(i - N / 2) / ((double)c * (double)N);
        phi[i] = orc_bessel_i0((double)K * sqrt(alpha * alpha - xi * xi));
    }
}

static double round_half_away(double v) { return v >= 0 ? floor(v + 0.5) : -floor(-v + 0.5); }
static double orc_round(double v, int mode) { return mode == 1 ? round_half_away(v) : nearbyint(v); }
static long pos_mod(long a, long n) { long r = a % n; return r < 0 ? r + n : r; }

/* -------------------------------------------------------------------------- */
/* Instantiate the transform bodies for double and float.                      */
#define REAL double
#define SFX(name) name##_d
#include "oracle_body.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f
#include "oracle_body.inc"
#undef REAL
#undef SFX

/* -------------------------------------------------------------------------- */
/* Direct sums (FP64 only).                                                     */

/* exp(-2 pi i t) with t reduced mod 1 first, so large phases keep full
 * precision. */
static void cis_neg(double t, double *re, double *im)
{
    t -= floor(t);
    double a = 2.0 * ORC_PI * t;
    *re = cos(a);
    *im = -sin(a);
}

/* NFFT/NDFT2_2D.cu:28-29,67: out[n] = sum_m in[m] exp(-2 pi i (u_n X_m / Nu + v_n Y_m / Nv)). */
void orc_ndft2_2d(const double *X, const double *Y, const double *u, const double *v,
                  const double *in, double *out, int Nu, int Nv, int M, int N)
{
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        double sr = 0, si = 0;
        for (int m = 0; m < M; ++m) {
            double t = u[n] * X[m] / (double)Nu + v[n] * Y[m] / (double)Nv;
            double cr, ci;
            cis_neg(t, &cr, &ci);
            sr += in[2 * m] * cr - in[2 * m + 1] * ci;
            si += in[2 * m] * ci + in[2 * m + 1] * cr;
        }
        out[2 * n] = sr;
        out[2 * n + 1] = si;
    }
}

/* Direct type-2 sum y_l = sum_{k1,k2} f[k1][k2] e^{-2 pi i (k1 x1_l / N1 + k2 x2_l / N2)},
 * k in [-N/2, N/2) at index k + N/2.  When dx1/dx2 are non-NULL also the
 * derivatives dy_l/dx1_l and dy_l/dx2_l. */
void orc_ndft_type2(double *y, double *dx1, double *dx2, const double *f,
                    const double *x1, const double *x2, int N1, int N2, int M)
{
#pragma omp parallel
    {
        double *e1 = (double *)malloc(sizeof(double) * 2 * N1);
        double *e2 = (double *)malloc(sizeof(double) * 2 * N2);
#pragma omp for schedule(static)
        for (int l = 0; l < M; ++l) {
            for (int i = 0; i < N1; ++i) cis_neg((double)(i - N1 / 2) * x1[l] / (double)N1, &e1[2 * i], &e1[2 * i + 1]);
            for (int j = 0; j < N2; ++j) cis_neg((double)(j - N2 / 2) * x2[l] / (double)N2, &e2[2 * j], &e2[2 * j + 1]);
            double yr = 0, yi = 0, ar = 0, ai = 0, br = 0, bi = 0;
            for (int i = 0; i < N1; ++i) {
                double sr = 0, si = 0, tr = 0, ti = 0;
                const double *row = f + (size_t)2 * i * N2;
                for (int j = 0; j < N2; ++j) {
                    double pr = row[2 * j] * e2[2 * j] - row[2 * j + 1] * e2[2 * j + 1];
                    double pi = row[2 * j] * e2[2 * j + 1] + row[2 * j + 1] * e2[2 * j];
                    sr += pr; si += pi;
                    double k2 = (double)(j - N2 / 2);
                    tr += k2 * pr; ti += k2 * pi;
                }
                double cr = e1[2 * i], ci = e1[2 * i + 1];
                double qr = sr * cr - si * ci, qi = sr * ci + si * cr;
                yr += qr; yi += qi;
                double k1 = (double)(i - N1 / 2);
                ar += k1 * qr; ai += k1 * qi;
                br += tr * cr - ti * ci; bi += tr * ci + ti * cr;
            }
            y[2 * l] = yr; y[2 * l + 1] = yi;
            /* d/dx e^{-2 pi i k x / N} = (-2 pi i k / N) e^{...} */
            if (dx1) { double s = -2.0 * ORC_PI / (double)N1; dx1[2 * l] = -s * ai; dx1[2 * l + 1] = s * ar; }
            if (dx2) { double s = -2.0 * ORC_PI / (double)N2; dx2[2 * l] = -s * bi; dx2[2 * l + 1] = s * br; }
        }
        free(e1);
        free(e2);
    }
}

/* Direct type-1 sum F[k1][k2] = sum_l d_l e^{-2 pi i (k1 x1_l / N1 + k2 x2_l / N2)}. */
void orc_ndft_type1(double *F, const double *d, const double *x1, const double *x2,
                    int N1, int N2, int M)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; ++i) {
        double *acc = (double *)calloc((size_t)2 * N2, sizeof(double));
        for (int l = 0; l < M; ++l) {
            double cr, ci;
            cis_neg((double)(i - N1 / 2) * x1[l] / (double)N1, &cr, &ci);
            double ar = d[2 * l] * cr - d[2 * l + 1] * ci, ai = d[2 * l] * ci + d[2 * l + 1] * cr;
            for (int j = 0; j < N2; ++j) {
                double er, ei;
                cis_neg((double)(j - N2 / 2) * x2[l] / (double)N2, &er, &ei);
                acc[2 * j] += ar * er - ai * ei;
                acc[2 * j + 1] += ar * ei + ai * er;
            }
        }
        memcpy(F + (size_t)2 * i * N2, acc, sizeof(double) * 2 * N2);
        free(acc);
    }
}

/* Direct type-1 sum at a subset of modes: F[q] = sum_l d_l e^{-2 pi i (k1_q x1_l / N1 + k2_q x2_l / N2)}
 * with (k1_q, k2_q) = (i1[q] - N1/2, i2[q] - N2/2). */
void orc_ndft_type1_modes(double *F, const double *d, const double *x1, const double *x2, int N1, int N2, int M,
                          const int *i1, const int *i2, int nq)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nq; ++q) {
        double k1 = (double)(i1[q] - N1 / 2), k2 = (double)(i2[q] - N2 / 2);
        double sr = 0, si = 0;
        for (int l = 0; l < M; ++l) {
            double cr, ci;
            cis_neg(k1 * x1[l] / (double)N1 + k2 * x2[l] / (double)N2, &cr, &ci);
            sr += d[2 * l] * cr - d[2 * l + 1] * ci;
            si += d[2 * l] * ci + d[2 * l + 1] * cr;
        }
        F[2 * q] = sr;
        F[2 * q + 1] = si;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* -------------------------------------------------------------------------- */
/* 1-D transforms (FP64), one line of a batch per call.                          */

/* Type-2, BP_FBP_FFBP/NFFT1_1D.m step by step (c = 2, K = 6 there):
 *   phi_k = I0(K sqrt(alpha^2 - xi_k^2)) / I0(K alpha), xi_k = 2 pi k / (c N)   :12-14
 *   mu_l = round(c x_l)                                                           :20
 *   spectrum_phi[l][m] = phi_hat(c x_l - (mu_l + m)) / I0(K alpha), m = -K..K   :22-30
 *   u = zeros(c N); u[(c-1) N/2 .. (c+1) N/2) = data ./ phi                      :34-35
 *   U = fft(ifftshift(u))                                                         :41
 *   result_l = sum_m spectrum_phi[l][m] U[(mu_l + m) mod c N]                     :47-48
 * data: N complex (mode k at index k + N/2), x: M grid-unit coordinates
 * (period N), result: M complex. */
void orc_nfft1_1d(double *result, const double *data, const double *x, int N, int M, int K, int c,
                  int round_mode)
{
    const double alpha = orc_alpha(c), S = orc_bessel_i0((double)K * alpha);
    const long L = (long)c * N;
    double *phi = (double *)malloc(sizeof(double) * N);
    double *u = (double *)calloc(2 * L, sizeof(double));
    double *U = (double *)malloc(sizeof(double) * 2 * L);
    double *buf = (double *)malloc(sizeof(double) * 2 * L), *tmp = (double *)malloc(sizeof(double) * 2 * L);
    orc_spectral_window(phi, N, K, c);
    for (int i = 0; i < N; ++i) {
        const long j = (long)(c - 1) * N / 2 + i;
        u[2 * j] = data[2 * i] / (phi[i] / S);
        u[2 * j + 1] = data[2 * i + 1] / (phi[i] / S);
    }
    /* ifftshift: U[n] = u[(n + L/2) mod L], then the forward DFT */
    for (long n = 0; n < L; ++n) {
        const long s = (n + L / 2) % L;
        U[2 * n] = u[2 * s];
        U[2 * n + 1] = u[2 * s + 1];
    }
    line_dft_d(U, L, 1, -1, buf, tmp);
#pragma omp parallel for schedule(static)
    for (int l = 0; l < M; ++l) {
        const double cx = (double)c * x[l], mu = orc_round(cx, round_mode);
        double sr = 0, si = 0;
        for (int m = -K; m <= K; ++m) {
            const double w = orc_window(cx - (mu + m), K, alpha) / S;
            const long n = pos_mod((long)mu + m, L);
            sr += w * U[2 * n];
            si += w * U[2 * n + 1];
        }
        result[2 * l] = sr;
        result[2 * l + 1] = si;
    }
    free(phi); free(u); free(U); free(buf); free(tmp);
}

/* Type-1, NFFT/NFFT2/1D/Matlab/NFFT2_1D.m step by step: M samples data_l at
 * x_l (grid units, period N) to N outputs:
 *   phi_k = I0(K sqrt(alpha^2 - xi_k^2)), xi_k = 2 pi k / (c N) (not normalised) :22-24
 *   u[k + c N/2] += data_l phi_hat(c x_l - k) for every k in [-c N/2, c N/2) with
 *       |round(c x_l) - k| <= K, or the same test one period c N up or down     :59-78
 *   FFT = fftshift(fft(ifftshift(u)))                                             :85
 *   result = FFT[(c-1) N/2 .. (c+1) N/2) ./ phi                                   :91 */
void orc_nfft2_1d(double *result, const double *data, const double *x, int N, int M, int K, int c,
                  int round_mode)
{
    const double alpha = orc_alpha(c);
    const long L = (long)c * N;
    double *phi = (double *)malloc(sizeof(double) * N);
    double *u = (double *)calloc(2 * L, sizeof(double));
    double *U = (double *)malloc(sizeof(double) * 2 * L);
    double *buf = (double *)malloc(sizeof(double) * 2 * L), *tmp = (double *)malloc(sizeof(double) * 2 * L);
    orc_spectral_window(phi, N, K, c);
    for (int l = 0; l < M; ++l) {
        const double cx = (double)c * x[l], r = orc_round(cx, round_mode);
        for (long k = -L / 2; k < L / 2; ++k) {
            double d = 0;
            int hit = 1;
            if (fabs(r - (double)k) <= K) d = cx - (double)k;
            else if (fabs(r + (double)L - (double)k) <= K) d = cx - (double)k + (double)L;
            else if (fabs(r - (double)L - (double)k) <= K) d = cx - (double)k - (double)L;
            else hit = 0;
            if (!hit) continue;
            const double w = orc_window(d, K, alpha);
            const long j = k + L / 2;
            u[2 * j] += data[2 * l] * w;
            u[2 * j + 1] += data[2 * l + 1] * w;
        }
    }
    for (long n = 0; n < L; ++n) {               /* ifftshift */
        const long s = (n + L / 2) % L;
        U[2 * n] = u[2 * s];
        U[2 * n + 1] = u[2 * s + 1];
    }
    line_dft_d(U, L, 1, -1, buf, tmp);
    for (int i = 0; i < N; ++i) {                /* fftshift, crop, deconvolve */
        const long j = (long)(c - 1) * N / 2 + i;      /* index in the shifted spectrum */
        const long n = (j + L / 2) % L;                /* its unshifted position */
        result[2 * i] = U[2 * n] / phi[i];
        result[2 * i + 1] = U[2 * n + 1] / phi[i];
    }
    free(phi); free(u); free(U); free(buf); free(tmp);
}

/* Direct 1-D sums, the closed forms of NFFT/NFFT1/1D/Matlab/NFFT1_1D_BLAS.m and
 * NFFT/NFFT2/1D/Matlab/NFFT2_1D_BLAS.m in grid units:
 *   type-2: y_l = sum_k f_k e^{-2 pi i k x_l / N}   (f: N, k at index k + N/2)
 *   type-1: F_k = sum_l d_l e^{-2 pi i k x_l / N}   (F: N)                     */
void orc_ndft1d_type2(double *y, const double *f, const double *x, int N, int M)
{
#pragma omp parallel for schedule(static)
    for (int l = 0; l < M; ++l) {
        double sr = 0, si = 0;
        for (int i = 0; i < N; ++i) {
            double cr, ci;
            cis_neg((double)(i - N / 2) * x[l] / (double)N, &cr, &ci);
            sr += f[2 * i] * cr - f[2 * i + 1] * ci;
            si += f[2 * i] * ci + f[2 * i + 1] * cr;
        }
        y[2 * l] = sr;
        y[2 * l + 1] = si;
    }
}

void orc_ndft1d_type1(double *F, const double *d, const double *x, int N, int M)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double sr = 0, si = 0;
        for (int l = 0; l < M; ++l) {
            double cr, ci;
            cis_neg((double)(i - N / 2) * x[l] / (double)N, &cr, &ci);
            sr += d[2 * l] * cr - d[2 * l + 1] * ci;
            si += d[2 * l] * ci + d[2 * l + 1] * cr;
        }
        F[2 * i] = sr;
        F[2 * i + 1] = si;
    }
}

/* -------------------------------------------------------------------------- */
/* Type-3 (non-uniform to non-uniform), FP64.                                    */

/* Parameters of NFFT/NFFT3/C/NFFT3_2D_C.cpp:214-267 (NUFFT_2D_CPU) from the data
 * extents: R = 2.2, msp (taps half-width before and after :267), the grid
 * Mrx x Mry, its steps Dx, Dy, the Gaussian's b and t1, the centres and the
 * rescaling kk_x, kk_y of x (divided) and s (multiplied). */
typedef struct {
    double xb, yb, sb, tb, kkx, kky, Dx, Dy, b, t1, fact;
    int msp, Mrx, Mry;
} orc_nufft3_par;

static void minmax(const double *a, int n, double *mx, double *mn)
{
    double hi = -(double)FLT_MAX, lo = (double)FLT_MAX;   /* the start values of :214-215 */
    for (int i = 0; i < n; ++i) {
        if (a[i] > hi) hi = a[i];
        if (a[i] < lo) lo = a[i];
    }
    *mx = hi;
    *mn = lo;
}

void orc_nufft3_params(const double *x, const double *y, const double *s, const double *t, double eps, int len_in,
                       int len_out, double *out /* 11 doubles + 3 ints as doubles */)
{
    double Mx, mx, My, my, Ms, ms, Mt, mt;
    minmax(x, len_in, &Mx, &mx);
    minmax(y, len_in, &My, &my);
    minmax(s, len_out, &Ms, &ms);
    minmax(t, len_out, &Mt, &mt);
    double xb = (Mx + mx) / 2, X1 = fabs(Mx - xb);
    double yb = (My + my) / 2, Y1 = fabs(My - yb);
    double sb = (Ms + ms) / 2, S = fabs(Ms - sb);
    double tb = (Mt + mt) / 2, T = fabs(Mt - tb);
    const double R = 2.2;
    double msp = 2 + 2 * (R * R) * (-log(eps / 76)) / (ORC_PI * (R * R - 2));
    int Mrx = 2 * (int)ceil((X1 * S / ORC_PI) * R * R + R * msp);
    int Mry = 2 * (int)ceil((Y1 * T / ORC_PI) * R * R + R * msp);
    double Dx = 2 * ORC_PI / Mrx, Dy = 2 * ORC_PI / Mry;
    double kkx = X1 / (ORC_PI / 2), kky = Y1 / (ORC_PI / 2);
    Mx /= kkx; mx /= kkx; xb = (Mx + mx) / 2;
    Ms *= kkx; ms *= kkx; sb = (Ms + ms) / 2;
    My /= kky; my /= kky; yb = (My + my) / 2;
    Mt *= kky; mt *= kky; tb = (Mt + mt) / 2;
    double b = (msp - 2) / (4 * ORC_PI), t1 = 1 / (4 * b);
    msp = floor(msp - 2);
    double fact = 1. / (4. * ORC_PI * b);
    out[0] = xb; out[1] = yb; out[2] = sb; out[3] = tb; out[4] = kkx; out[5] = kky;
    out[6] = Dx; out[7] = Dy; out[8] = b; out[9] = t1; out[10] = fact * fact;
    out[11] = msp; out[12] = Mrx; out[13] = Mry;
}

/* NUFFT_2D_CPU step by step: spatialConvolutionFunction (:38-85, Gaussian
 * gridding of the pre-modulated samples), the forward 2-D DFT of the
 * Mry x Mrx grid (:287-292), circAndFFTShiftsFunction (:135-151, half-size
 * shift with the (-1)^(i+j) checkerboard), and
 * spectralConvolutionFunctionCompensation (:156-194, Gaussian interpolation
 * at the outputs and the post-compensation).  f, F interleaved complex.
 * Returns the grid size in grid[0..1] (Mrx, Mry).  mutation (tests only):
 * 1 drops the checkerboard sign, 2 the post-compensation factor, 3 flips the
 * sign of the pre-modulation phase. */
void orc_nufft3_2d(double *F, const double *x0, const double *y0, const double *s0, const double *t0, const double *f,
                   double eps, int len_in, int len_out, int *grid, int mutation)
{
    double par[14];
    orc_nufft3_params(x0, y0, s0, t0, eps, len_in, len_out, par);
    const double xb = par[0], yb = par[1], sb = par[2], tb = par[3], kkx = par[4], kky = par[5];
    const double Dx = par[6], Dy = par[7], b = par[8], t1 = par[9], fact = par[10];
    const int msp = (int)par[11], Mrx = (int)par[12], Mry = (int)par[13];
    grid[0] = Mrx;
    grid[1] = Mry;
    double *g = (double *)calloc((size_t)2 * Mrx * Mry, sizeof(double));
    for (int k = 0; k < len_in; ++k) {
        const double x = x0[k] / kkx, y = y0[k] / kky;
        const double arg = (mutation == 3 ? -1 : 1) * (sb * x + tb * y);
        const double re = f[2 * k] * cos(arg) + f[2 * k + 1] * sin(arg);
        const double im = f[2 * k + 1] * cos(arg) - f[2 * k] * sin(arg);
        const double xt = (x - xb) / Dx, yt = (y - yb) / Dy;
        const int nno = (int)floor(Mrx / 2 + xt), mmo = (int)floor(Mry / 2 + yt);
        for (int mm = -msp; mm <= msp; ++mm) {
            const int j = mm + mmo;
            const double dsj = j - Mry / 2;
            for (int nn = -msp; nn <= msp; ++nn) {
                const int i = nn + nno;
                const double dsi = i - Mrx / 2;
                const double e = fact * exp(b * (Dy * Dy * dsj * dsj + Dx * Dx * dsi * dsi) -
                                            t1 * ((xt - dsi) * (xt - dsi) + (yt - dsj) * (yt - dsj)));
                const size_t idx = (size_t)j * Mrx + i;
                g[2 * idx] += re * e;
                g[2 * idx + 1] += im * e;
            }
        }
    }
    /* forward 2-D DFT: rows (length Mrx), then columns (length Mry) */
    {
        const int nmax = Mrx > Mry ? Mrx : Mry;
        double *buf = (double *)malloc(sizeof(double) * 2 * nmax), *tmp = (double *)malloc(sizeof(double) * 2 * nmax);
        for (int j = 0; j < Mry; ++j) line_dft_d(g + (size_t)2 * j * Mrx, Mrx, 1, -1, buf, tmp);
        for (int i = 0; i < Mrx; ++i) line_dft_d(g + (size_t)2 * i, Mry, Mrx, -1, buf, tmp);
        free(buf); free(tmp);
    }
    /* shift with the checkerboard sign */
    double *gs = (double *)malloc(sizeof(double) * 2 * (size_t)Mrx * Mry);
    for (int i = 0; i < Mrx; ++i)
        for (int j = 0; j < Mry; ++j) {
            const int ii = (i + Mrx / 2) % Mrx, jj = (j + Mry / 2) % Mry;
            const double sg = mutation == 1 ? 1. : 1. - 2 * ((i + j) & 1);
            gs[2 * ((size_t)jj * Mrx + ii)] = sg * g[2 * ((size_t)j * Mrx + i)];
            gs[2 * ((size_t)jj * Mrx + ii) + 1] = sg * g[2 * ((size_t)j * Mrx + i) + 1];
        }
    for (int kk = 0; kk < len_out; ++kk) {
        const double s = s0[kk] * kkx, t = t0[kk] * kky;
        const int nsn = (int)floor(Mrx / 2 + (s - sb));
        const double difs = (Mrx / 2 + (s - sb)) - nsn;
        const int ntn = (int)floor(Mry / 2 + (t - tb));
        const double dift = (Mry / 2 + (t - tb)) - ntn;
        const double temp = mutation == 2 ? 1. : exp(b * (Dy * (t - tb) * Dy * (t - tb) + Dx * (s - sb) * Dx * (s - sb)));
        double ar = 0, ai = 0;
        for (int nn = nsn - msp; nn <= nsn + msp; ++nn)
            for (int mm = ntn - msp; mm <= ntn + msp; ++mm) {
                const double npi = mm - ntn, npj = nn - nsn;
                const double e = exp(-t1 * ((dift - npi) * (dift - npi) + (difs - npj) * (difs - npj)));
                const size_t idx = (size_t)mm * Mrx + nn;
                ar += gs[2 * idx] * e;
                ai += gs[2 * idx + 1] * e;
            }
        const double arg = (s - sb) * xb + (t - tb) * yb;
        F[2 * kk] = (ar * cos(arg) + ai * sin(arg)) * temp;
        F[2 * kk + 1] = (ai * cos(arg) - ar * sin(arg)) * temp;
    }
    free(g);
    free(gs);
}

/* Direct type-3 sum F_j = sum_k f_k e^{-i (s_j x_k + t_j y_k)} (the quantity
 * NUFFT_2D_CPU approximates; phases are not reduced: cos/sin of FP64 take any
 * argument). */
void orc_ndft3_2d(double *F, const double *x, const double *y, const double *s, const double *t, const double *f,
                  int len_in, int len_out)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < len_out; ++j) {
        double ar = 0, ai = 0;
        for (int k = 0; k < len_in; ++k) {
            const double a = s[j] * x[k] + t[j] * y[k], c = cos(a), sn = sin(a);
            ar += f[2 * k] * c + f[2 * k + 1] * sn;
            ai += f[2 * k + 1] * c - f[2 * k] * sn;
        }
        F[2 * j] = ar;
        F[2 * j + 1] = ai;
    }
}
