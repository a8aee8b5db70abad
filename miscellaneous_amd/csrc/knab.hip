// knab.hip -- the element-wise steps of the band-limited polar interpolation
// (BP_FBP_FFBP/knabNFFT2D.m, rectNFFT2D.m) around its type-2 NFFT:
//   * the grid steps, origins and extents that size the transform
//     (knabNFFT2D.m:28-40), one pass over the output points (partial minima
//     and maxima per block, folded on the host with the one transfer the
//     sizing needs anyway);
//   * the spectrum of the samples repeated periodically, windowed, zero padded
//     to a N x b M and conjugated (knabNFFT2D.m:58-70 with
//     NFFT_Based_Selvas_Approach_for_Signal_Interp_2D.m:9-20), one pass;
//   * the output points in units of the padded grid, wrapped to one period in
//     FP64 and rounded once to the transform's type (knabNFFT2D.m:72-73), one pass;
//   * the conjugated, scaled and cropped result with ExtrapVal outside the
//     input grid (knabNFFT2D.m:74-79), one pass.
// The 2-D FFT of the samples and the type-2 NFFT in between are separate calls.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/nfft2d.h"
#include "nfft2d_device.hpp"
#include "nfft2d_internal.hpp"

namespace knab {

using nfft2d::cplx;

constexpr int kThreads = 256;
constexpr int kStatBlocks = NFFT2D_KNAB_STAT_BLOCKS;

// NaN-propagating min / max (a NaN extent sends the caller to its Nn = 1 branch)
__device__ __forceinline__ double nmin(double a, double b) { return (a != a || a < b) ? a : b; }
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || a > b) ? a : b; }

// Wave-then-block reduction over kThreads threads (4 waves); every thread gets the result.
template <typename F>
__device__ __forceinline__ double block_reduce(double v, double* sh, F op)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();                    // sh free from any earlier call
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    return op(op(sh[0], sh[1]), op(sh[2], sh[3]));
}

__device__ __forceinline__ double dsum(double a, double b) { return a + b; }

// Element strides of the output coordinate matrices (rho_out row and column,
// cos_out row and column): meshgrid views pass in without a copy.
struct Strides {
    long long rr, rc, cr, cc;
};

// Block b: min and max of ro and co over its grid-stride share ->
// part[4 b .. 4 b + 3].  Block 0 also writes the axis figures to out[0..9]
// (out[4], out[5] are left to the host, which folds the partials).
__global__ __launch_bounds__(kThreads) void k_stats(const double* __restrict__ rax, int nr, long long rs,
                                                    const double* __restrict__ cax, int nc, long long cs,
                                                    const double* __restrict__ ro, const double* __restrict__ co,
                                                    int R, int C, Strides st, double* __restrict__ out,
                                                    double* __restrict__ part)
{
    __shared__ double sh[kThreads / 64];
    const double inf = __builtin_huge_val();
    double a0 = inf, a1 = -inf, b0 = inf, b1 = -inf;
    const long long n = (long long)R * C;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const long long r = i / C, c = i % C;
        const double rv = ro[r * st.rr + c * st.rc], cv = co[r * st.cr + c * st.cc];
        a0 = nmin(a0, rv);
        a1 = nmax(a1, rv);
        b0 = nmin(b0, cv);
        b1 = nmax(b1, cv);
    }
    a0 = block_reduce(a0, sh, nmin);
    a1 = block_reduce(a1, sh, nmax);
    b0 = block_reduce(b0, sh, nmin);
    b1 = block_reduce(b1, sh, nmax);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = a0;
        part[4 * blockIdx.x + 1] = a1;
        part[4 * blockIdx.x + 2] = b0;
        part[4 * blockIdx.x + 3] = b1;
    }
    if (blockIdx.x != 0) return;
    const int t = threadIdx.x;
    double dr = 0, dc = 0, r0 = inf, r1 = -inf, c0 = inf, c1 = -inf;
    for (int k = t; k < nr; k += kThreads) {
        const double v = rax[k * rs];
        if (k + 1 < nr) dr += rax[(k + 1) * rs] - v;
        r0 = nmin(r0, v);
        r1 = nmax(r1, v);
    }
    for (int k = t; k < nc; k += kThreads) {
        const double v = cax[k * cs];
        if (k + 1 < nc) dc += cax[(k + 1) * cs] - v;
        c0 = nmin(c0, v);
        c1 = nmax(c1, v);
    }
    dr = block_reduce(dr, sh, dsum);
    dc = block_reduce(dc, sh, dsum);
    r0 = block_reduce(r0, sh, nmin);
    r1 = block_reduce(r1, sh, nmax);
    c0 = block_reduce(c0, sh, nmin);
    c1 = block_reduce(c1, sh, nmax);
    if (t == 0) {
        out[0] = dr / (nr - 1);
        out[1] = dc / (nc - 1);
        out[2] = rax[0];
        out[3] = cax[0];
        out[4] = 0.0;
        out[5] = 0.0;
        out[6] = r0;
        out[7] = r1;
        out[8] = c0;
        out[9] = c1;
    }
}

// F[p][q], p < PN, q < QM: the band |p - PN/2| <= kgn, |q - QM/2| <= kgm holds
// conj(S[(v - kgm) mod M][(u - kgn) mod N] w(u, v)), u = p - PN/2 + kgn,
// v = q - QM/2 + kgm; zero elsewhere.  S is the caller's M x N spectrum
// (cos alpha x rho), read transposed; w = Gn[u] Gm[v] scale in FP64, rounded
// once to T, or the scale alone without tables.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spectrum(const cplx<T>* __restrict__ S, int M, int N, int kgn, int kgm,
                                                       const double* __restrict__ Gn, const double* __restrict__ Gm,
                                                       double scale, cplx<T>* __restrict__ F, int PN, int QM)
{
    const int p = blockIdx.y;
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= QM) return;
    const int u = p - PN / 2 + kgn, v = q - QM / 2 + kgm;
    cplx<T> o;
    o.x = (T)0;
    o.y = (T)0;
    if (u >= 0 && u <= 2 * kgn && v >= 0 && v <= 2 * kgm) {
        int i = (u - kgn) % N, j = (v - kgm) % M;
        i += i < 0 ? N : 0;
        j += j < 0 ? M : 0;
        const cplx<T> s = S[(size_t)j * N + i];
        const T w = Gn ? (T)(Gn[u] * Gm[v] * scale) : (T)scale;
        o.x = s.x * w;
        o.y = -(s.y * w);
    }
    F[(size_t)p * QM + q] = o;
}

// x1[r][c] = wrap((ro[r][c] - r0) / h1, P1), x2 likewise, over the padded
// R2 x C2 layout; entries outside the caller's R x C take 0.
// wrap(x, P) = x - P rint(x / P), every step rounded in FP64 as written.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_points(const double* __restrict__ ro, const double* __restrict__ co,
                                                     int R, int C, Strides st, int R2, int C2, double r0, double c0,
                                                     double h1, double h2, double P1, double P2, T* __restrict__ x1,
                                                     T* __restrict__ x2)
{
#pragma clang fp contract(off)
    const long long n = (long long)R2 * C2;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const long long r = i / C2, c = i % C2;
        double a = 0.0, b = 0.0;
        if (r < R && c < C) {
            a = (ro[r * st.rr + c * st.rc] - r0) / h1;
            b = (co[r * st.cr + c * st.cc] - c0) / h2;
        }
        a = a - P1 * rint(a / P1);
        b = b - P2 * rint(b / P2);
        x1[i] = (T)a;
        x2[i] = (T)b;
    }
}

// res[r][c] = conj(y[r C2 + c]) / s, or ExtrapVal where ro[0][c] or co[r][0]
// lies outside the input grid's range.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_finish(const cplx<T>* __restrict__ y, int R, int C, int C2, double s,
                                                     const double* __restrict__ ro, const double* __restrict__ co,
                                                     Strides st, double rmin, double rmax, double cmin, double cmax,
                                                     double ex_re, double ex_im, cplx<T>* __restrict__ res)
{
    const long long n = (long long)R * C;
    const T sc = (T)s;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const long long r = i / C, c = i % C;
        const double rv = ro[c * st.rc], cv = co[r * st.cr];
        cplx<T> o;
        if (rv < rmin || rv > rmax || cv < cmin || cv > cmax) {
            o.x = (T)ex_re;
            o.y = (T)ex_im;
        } else {
            const cplx<T> v = y[r * C2 + c];
            o.x = v.x / sc;
            o.y = -v.y / sc;
        }
        res[i] = o;
    }
}

inline int grid_for(long long n)
{
    const long long b = (n + kThreads - 1) / kThreads;
    return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

inline Strides strides_of(const long long *s, int C)
{
    return s ? Strides{s[0], s[1], s[2], s[3]} : Strides{C, 1, C, 1};
}

}  // namespace knab

using namespace knab;

extern "C" {

int nfft2d_knab_stats(const double *rho_axis, int n_rho, long long rho_stride, const double *cos_axis, int n_cos,
                      long long cos_stride, const double *rho_out, const double *cos_out, int R, int C,
                      const long long *out_strides, double *out, void *stream)
{
    if (!rho_axis || !cos_axis || !rho_out || !cos_out || !out)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_stats: null pointer");
    if (n_rho < 2 || n_cos < 2 || R < 1 || C < 1 || rho_stride < 1 || cos_stride < 1)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_stats: axes need 2 or more samples and R, C >= 1 (got %d, %d, %d, %d)",
                           n_rho, n_cos, R, C);
    // every block writes its partial (+-inf when its share is empty)
    hipLaunchKernelGGL(k_stats, dim3(kStatBlocks), dim3(kThreads), 0, (hipStream_t)stream, rho_axis, n_rho, rho_stride,
                       cos_axis, n_cos, cos_stride, rho_out, cos_out, R, C, strides_of(out_strides, C), out,
                       out + NFFT2D_KNAB_STATS);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

int nfft2d_knab_spectrum(int dtype, const void *S, int M, int N, int kgn, int kgm, const double *Gn, const double *Gm,
                         double scale, void *F, int PN, int QM, void *stream)
{
    if (!S || !F || (!Gn) != (!Gm)) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_spectrum: null pointer");
    if (dtype != NFFT2D_C64 && dtype != NFFT2D_C128)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_spectrum: dtype %d", dtype);
    if (M < 1 || N < 1 || kgn < 0 || kgm < 0 || PN < 2 * kgn + 1 || QM < 2 * kgm + 1 || PN > 65535)
        return nfft2d_fail(NFFT2D_ERR_ARG,
                           "nfft2d_knab_spectrum: the band 2 kg + 1 (%d, %d) must fit the padded %d x %d (rows <= 65535)",
                           2 * kgn + 1, 2 * kgm + 1, PN, QM);
    const dim3 grid((QM + kThreads - 1) / kThreads, PN);
    const hipStream_t st = (hipStream_t)stream;
    if (dtype == NFFT2D_C64)
        hipLaunchKernelGGL(k_spectrum<float>, grid, dim3(kThreads), 0, st, (const float2 *)S, M, N, kgn, kgm, Gn, Gm,
                           scale, (float2 *)F, PN, QM);
    else
        hipLaunchKernelGGL(k_spectrum<double>, grid, dim3(kThreads), 0, st, (const double2 *)S, M, N, kgn, kgm, Gn, Gm,
                           scale, (double2 *)F, PN, QM);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

int nfft2d_knab_points(int dtype, const double *rho_out, const double *cos_out, int R, int C,
                       const long long *out_strides, int R2, int C2, double r0, double c0, double h1, double h2,
                       double P1, double P2, void *x1, void *x2, void *stream)
{
    if (!rho_out || !cos_out || !x1 || !x2) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_points: null pointer");
    if (dtype != NFFT2D_C64 && dtype != NFFT2D_C128)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_points: dtype %d", dtype);
    if (R < 1 || C < 1 || R2 < R || C2 < C || !(h1 != 0) || !(h2 != 0) || !(P1 > 0) || !(P2 > 0))
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_points: bad sizes or steps (%d x %d in %d x %d)", R, C, R2, C2);
    const int g = grid_for((long long)R2 * C2);
    const hipStream_t st = (hipStream_t)stream;
    const Strides s = strides_of(out_strides, C);
    if (dtype == NFFT2D_C64)
        hipLaunchKernelGGL(k_points<float>, dim3(g), dim3(kThreads), 0, st, rho_out, cos_out, R, C, s, R2, C2, r0, c0,
                           h1, h2, P1, P2, (float *)x1, (float *)x2);
    else
        hipLaunchKernelGGL(k_points<double>, dim3(g), dim3(kThreads), 0, st, rho_out, cos_out, R, C, s, R2, C2, r0, c0,
                           h1, h2, P1, P2, (double *)x1, (double *)x2);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

int nfft2d_knab_finish(int dtype, const void *y, int R, int C, int C2, double scale, const double *rho_out,
                       const double *cos_out, const long long *out_strides, double rho_min, double rho_max,
                       double cos_min, double cos_max, double extrap_re, double extrap_im, void *result, void *stream)
{
    if (!y || !rho_out || !cos_out || !result) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_finish: null pointer");
    if (dtype != NFFT2D_C64 && dtype != NFFT2D_C128)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_finish: dtype %d", dtype);
    if (R < 1 || C < 1 || C2 < C) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_knab_finish: bad sizes");
    const int g = grid_for((long long)R * C);
    const hipStream_t st = (hipStream_t)stream;
    const Strides s = strides_of(out_strides, C);
    if (dtype == NFFT2D_C64)
        hipLaunchKernelGGL(k_finish<float>, dim3(g), dim3(kThreads), 0, st, (const float2 *)y, R, C, C2, scale, rho_out,
                           cos_out, s, rho_min, rho_max, cos_min, cos_max, extrap_re, extrap_im, (float2 *)result);
    else
        hipLaunchKernelGGL(k_finish<double>, dim3(g), dim3(kThreads), 0, st, (const double2 *)y, R, C, C2, scale,
                           rho_out, cos_out, s, rho_min, rho_max, cos_min, cos_max, extrap_re, extrap_im,
                           (double2 *)result);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

}  // extern "C"
