// ndft2d.hip -- the 2D NDFT.  Part 1: NDFT2_2D_GPU, the reference's direct
// sum in FP64 (NFFT/NDFT2_2D.cu:15-32,37-72):
//   out[n] = sum_{m<M} in[m] exp(-2 pi i (u[n] X[m] / Nu + v[n] Y[m] / Nv)).
// The reference fills an explicit M x N FP64 kernel matrix (:44-50; 64 GiB at
// M = N = 65536), writes it to a text file (:56) and calls cublasZgemv (:67).
// Here nothing of that size is materialised: on the integer meshgrid the
// reference's caller passes, the sum is an FP64 complex GEMM on the matrix
// cores (k_ndft_grid + k_ndft_grid_sum, with a cached workspace of at most 64
// partial output arrays); for other (u, v) each thread owns one output n and
// streams the points through LDS in chunks, accumulating in FP64
// (k_ndft_direct).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>

#include "../../include/nfft2d.h"
#include "ndft2d_mfma.hpp"
#include "nfft2d_internal.hpp"

namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// Output grids.  The reference's only caller passes (u, v) = the integer
// meshgrid of the far-field plane (Synthesis/PSO/Synthesis.cu:85-90,397):
// N = nu nv outputs with u and v arithmetic progressions.  Then
//     out[i, j] = sum_m in_m e^{-2 pi i a_i X_m / Nu} e^{-2 pi i b_j Y_m / Nv},
//     a_i = u0 + i du, b_j = v0 + j dv,
// is the complex product (nu x M) . (M x nv) of two phasor matrices: a GEMM on
// the FP64 matrix cores instead of M N separate kernel evaluations.
// k_grid_detect recognises the two flattenings of such a grid (u constant
// over runs of nv outputs, or v constant over runs of nu) and k_grid_verify
// checks every output against it by exact FP64 comparison; the finding stays
// in g_grid (one per device, as every __device__ variable); k_ndft_grid runs
// when it is a grid, k_ndft_direct
// (any u, v) when it did not.  Both read the flag on the device, so the call
// stays asynchronous and capturable.  Two NDFT2_2D_GPU calls that overlap on
// one device (different streams) race on g_grid.
struct GridInfo {
    int ok;          // 1: (u, v) is a grid of nu x nv outputs
    int vmajor;      // 0: n = i nv + j (u constant over runs of nv); 1: n = j nu + i
    int nu, nv;
    double u0, du, v0, dv;
};
__device__ GridInfo g_grid;

constexpr int kDetThreads = 1024;
__global__ __launch_bounds__(kDetThreads) void k_grid_detect(const double *__restrict__ u,
                                                             const double *__restrict__ v, int N)
{
    __shared__ int s_run[2];
    const int t = threadIdx.x;
    if (t == 0) {
        s_run[0] = N;
        s_run[1] = N;
    }
    __syncthreads();
    const double u0 = u[0], v0 = v[0];
    // first change of u and of v, block-wide over chunks of kDetThreads
    // outputs: each wave's first hit is the lowest set bit of its ballot, one
    // LDS atomic per wave and value takes the block's minimum, and the block
    // stops at the first chunk after which both are found (later chunks hold
    // only larger indices).  An atomic per element serialised on one address
    // (425 us at 65,536 outputs); threads scanning on their own until their
    // own first hits ran the whole array where a stride of kDetThreads never
    // meets a change (29 us on a 256 x 256 grid).
    const int lane = t & 63;
    for (int base = 1; base < N; base += kDetThreads) {
        const int n = base + t;
        const bool in = n < N;
        const unsigned long long bu = __ballot(in && u[n] != u0);
        const unsigned long long bv = __ballot(in && v[n] != v0);
        if (lane == 0) {   // n is the wave's first index
            if (bu) atomicMin(&s_run[0], n + __ffsll((long long)bu) - 1);
            if (bv) atomicMin(&s_run[1], n + __ffsll((long long)bv) - 1);
        }
        __syncthreads();
        const bool done = s_run[0] < N && s_run[1] < N;
        __syncthreads();   // every thread has read before the next chunk's atomics
        if (done) break;
    }
    const int ru = s_run[0], rv = s_run[1];
    // u-major: v changes at once and u after nv = ru outputs; v-major: the converse
    int vmajor = -1, nu = 1, nv = 1;
    if (N == 1) {
        vmajor = 0;
    } else if (rv == 1 && ru > 1 && N % ru == 0) {
        vmajor = 0;
        nv = ru;
        nu = N / ru;
    } else if (ru == 1 && rv > 1 && N % rv == 0) {
        vmajor = 1;
        nu = rv;
        nv = N / rv;
    }
    double du = 0.0, dv = 0.0;
    if (vmajor == 0) {
        du = nu > 1 ? u[nv] - u0 : 0.0;
        dv = nv > 1 ? v[1] - v0 : 0.0;
    } else if (vmajor == 1) {
        du = nu > 1 ? u[1] - u0 : 0.0;
        dv = nv > 1 ? v[nu] - v0 : 0.0;
    }
    // every output is then checked against the grid by k_grid_verify (all of
    // the chip: one block's loop over N outputs took 76 us at N = 65,536)
    if (t == 0) {
        GridInfo g;
        g.ok = vmajor >= 0 && nu * nv == N;
        g.vmajor = vmajor > 0;
        g.nu = nu;
        g.nv = nv;
        g.u0 = u0;
        g.du = du;
        g.v0 = v0;
        g.dv = dv;
        g_grid = g;
    }
}

// Clears g_grid.ok when an output is not the grid's: u[n] = u0 + i du and
// v[n] = v0 + j dv, exactly, for n's (i, j) in the detected flattening.
__global__ __launch_bounds__(kThreads) void k_grid_verify(const double *__restrict__ u, const double *__restrict__ v,
                                                          int N)
{
    const GridInfo g = g_grid;
    if (!g.ok) return;
    const int n = blockIdx.x * kThreads + threadIdx.x;
    if (n >= N) return;
    const int i = g.vmajor ? n % g.nu : n / g.nv, j = g.vmajor ? n / g.nu : n % g.nv;
    if (u[n] != g.u0 + i * g.du || v[n] != g.v0 + j * g.dv) g_grid.ok = 0;
}

// The grid case on v_mfma_f64_16x16x4f64 (4 points per instruction): lane l
// supplies A[i][k] at i = l % 16, k = l / 16 and B[k][j] at j = l % 16;
// accumulator element r of lane l is C[4 r + l / 16][l % 16].  The complex
// product is four real ones: Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br.
//
// A work item is one wave: a supertile of kTR x kTC MFMA tiles (32 x 64
// outputs) over one range of the points.  Per K-step of 4 points a lane
// evaluates its point's phasor at its first row and first column by an FP64
// phase reduction and sincospi, and the phasors 16 rows / columns further on
// by one complex product with the point's 16-step phasor each (rows and
// columns are arithmetic progressions; the steps come from k_ndft_steps, once
// per call), so that 8 MFMA tiles share 2 sincospi per lane instead of 16.  The partial supertile goes to range r's slice of a
// workspace; k_ndft_grid_sum adds the ranges in a fixed order (the result does
// not depend on the launch).  Work items are walked persistently (their count
// is known on the device only).
constexpr int kTR = 2, kTC = 4;                   // MFMA tiles per supertile: rows, columns
constexpr int kSupRows = 16 * kTR, kSupCols = 16 * kTC;
constexpr int kGridWavesTotal = 2048;             // waves of the persistent launch (2 per SIMD at 190 VGPRs)
constexpr int kMaxRanges = 64;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 phasor(double t)   // e^{-2 pi i t}
{
    t -= floor(t);
    double s, c;
    sincospi(2.0 * t, &s, &c);
    return make_double2(c, -s);
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Supertiles per grid (device side, from the detector's nu, nv).
__device__ __forceinline__ int n_super(const GridInfo &g, int &stv)
{
    stv = (g.nv + kSupCols - 1) / kSupCols;
    return ((g.nu + kSupRows - 1) / kSupRows) * stv;
}

// steps[m] = (e^{-2 pi i 16 du X_m / Nu}, e^{-2 pi i 16 dv Y_m / Nv}): the 16-row and 16-column
// steps of point m's phasors, once per call instead of once per supertile.
__global__ __launch_bounds__(kThreads) void k_ndft_steps(const double *__restrict__ X, const double *__restrict__ Y,
                                                         double inv_Nu, double inv_Nv, int M,
                                                         double4 *__restrict__ steps)
{
    const GridInfo g = g_grid;
    if (!g.ok) return;
    const int m = blockIdx.x * kThreads + threadIdx.x;
    if (m >= M) return;
    const double2 sa = phasor(16.0 * g.du * (X[m] * inv_Nu)), sb = phasor(16.0 * g.dv * (Y[m] * inv_Nv));
    steps[m] = make_double4(sa.x, sa.y, sb.x, sb.y);
}

__global__ __launch_bounds__(256) void k_ndft_grid(const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double2 *__restrict__ in, const double4 *__restrict__ steps,
                                                   double2 *__restrict__ part, double inv_Nu, double inv_Nv, int M,
                                                   int ranges, int chunk)
{
    const GridInfo g = g_grid;
    if (!g.ok) return;
    const int lane = threadIdx.x & 63, q = lane & 15, k = lane >> 4;
    int stv;
    const int items = n_super(g, stv) * ranges;
    const long long N = (long long)g.nu * g.nv;
    for (int item = blockIdx.x * 4 + (threadIdx.x >> 6); item < items; item += gridDim.x * 4) {
        const int r = item % ranges, st = item / ranges;
        const int i0 = (st / stv) * kSupRows, j0 = (st % stv) * kSupCols;
        const int m_lo = min(M, r * chunk), m_hi = min(M, m_lo + chunk);
        const double a0 = g.u0 + (i0 + q) * g.du, b0 = g.v0 + (j0 + q) * g.dv;   // this lane's first row / column
        bool row_in[kTR], col_in[kTC];
#pragma unroll
        for (int t = 0; t < kTR; ++t) row_in[t] = i0 + 16 * t + q < g.nu;
#pragma unroll
        for (int t = 0; t < kTC; ++t) col_in[t] = j0 + 16 * t + q < g.nv;
        d4 cr[kTR][kTC], ci[kTR][kTC];
#pragma unroll
        for (int a = 0; a < kTR; ++a)
#pragma unroll
            for (int b = 0; b < kTC; ++b) {
                cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
                ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
            }
        // the point data of the next K-step is loaded while the current one is
        // multiplied (one load round trip per K-step would set the pace)
        struct Pt {
            double2 d;
            double x, y;
            double4 st;
        };
        auto load_pt = [&](int m) {
            Pt p;
            if (m < m_hi) {
                p.d = in[m];
                p.x = X[m];
                p.y = Y[m];
                p.st = steps[m];
            } else {
                p.d = make_double2(0.0, 0.0);
                p.x = p.y = 0.0;
                p.st = make_double4(1.0, 0.0, 1.0, 0.0);
            }
            return p;
        };
        Pt cur = load_pt(m_lo + k);
        for (int m0 = m_lo; m0 < m_hi; m0 += 4) {
            const Pt nxt = load_pt(m0 + 4 + k);
            double ar[kTR], ai[kTR], br[kTC], bi[kTC];
            {
                // a point past the range has d = 0: its A operands vanish
                const double xm = cur.x * inv_Nu, ym = cur.y * inv_Nv;
                double2 pa = phasor(a0 * xm), pb = phasor(b0 * ym);
                const double2 sa = make_double2(cur.st.x, cur.st.y), sb = make_double2(cur.st.z, cur.st.w);
#pragma unroll
                for (int t = 0; t < kTR; ++t) {
                    const double2 v = zmul(cur.d, pa);
                    ar[t] = row_in[t] ? v.x : 0.0;
                    ai[t] = row_in[t] ? v.y : 0.0;
                    if (t + 1 < kTR) pa = zmul(pa, sa);
                }
#pragma unroll
                for (int t = 0; t < kTC; ++t) {
                    br[t] = col_in[t] ? pb.x : 0.0;
                    bi[t] = col_in[t] ? pb.y : 0.0;
                    if (t + 1 < kTC) pb = zmul(pb, sb);
                }
            }
#pragma unroll
            for (int a = 0; a < kTR; ++a)
#pragma unroll
                for (int b = 0; b < kTC; ++b) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
                }
#pragma unroll
            for (int a = 0; a < kTR; ++a)
#pragma unroll
                for (int b = 0; b < kTC; ++b) {
                    cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], -bi[b], cr[a][b], 0, 0, 0);
                    ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
                }
            cur = nxt;
        }
        double2 *dst = part + (size_t)r * N;
#pragma unroll
        for (int a = 0; a < kTR; ++a)
#pragma unroll
            for (int b = 0; b < kTC; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = i0 + 16 * a + 4 * e + k, j = j0 + 16 * b + q;
                    if (i < g.nu && j < g.nv) dst[(size_t)i * g.nv + j] = make_double2(cr[a][b][e], ci[a][b][e]);
                }
    }
}

// out[n] = sum over the ranges (in order) of the partial sums, in the caller's flattening.
__global__ __launch_bounds__(kThreads) void k_ndft_grid_sum(const double2 *__restrict__ part, int ranges,
                                                            double2 *__restrict__ out, int N)
{
    const GridInfo g = g_grid;
    if (!g.ok) return;
    const int n = blockIdx.x * kThreads + threadIdx.x;
    if (n >= N) return;
    const int i = g.vmajor ? n % g.nu : n / g.nv, j = g.vmajor ? n / g.nu : n % g.nv;
    const size_t e = (size_t)i * g.nv + j;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < ranges; ++r) {
        const double2 v = part[(size_t)r * N + e];
        sr += v.x;
        si += v.y;
    }
    out[n] = make_double2(sr, si);
}

// The partial sums' workspace, one per device, grown as needed (the
// drop-in's calls run on the null stream, so consecutive calls are ordered).
struct GridWs {
    double2 *ptr = nullptr;
    size_t count = 0;
};
std::mutex g_ws_mu;
std::map<int, GridWs> g_ws;

__global__ __launch_bounds__(kThreads) void k_ndft_direct(const double *__restrict__ X, const double *__restrict__ Y,
                                                          const double *__restrict__ u, const double *__restrict__ v,
                                                          const double2 *__restrict__ in, double2 *__restrict__ out,
                                                          double inv_Nu, double inv_Nv, int M, int N)
{
    if (g_grid.ok) return;         // k_ndft_grid wrote the outputs
    __shared__ double sX[kThreads], sY[kThreads];
    __shared__ double2 sIn[kThreads];
    const int n = blockIdx.x * kThreads + threadIdx.x;
    const double un = n < N ? u[n] * inv_Nu : 0.0;
    const double vn = n < N ? v[n] * inv_Nv : 0.0;
    double ar = 0.0, ai = 0.0;
    for (int m0 = 0; m0 < M; m0 += kThreads) {
        const int m = m0 + threadIdx.x;
        if (m < M) {
            sX[threadIdx.x] = X[m];
            sY[threadIdx.x] = Y[m];
            sIn[threadIdx.x] = in[m];
        }
        __syncthreads();
        const int cnt = min(kThreads, M - m0);
        for (int q = 0; q < cnt; ++q) {
            double t = un * sX[q] + vn * sY[q];
            t -= floor(t);
            double s, c;
            sincospi(2.0 * t, &s, &c);   // e^{-2 pi i t} = c - i s
            const double2 d = sIn[q];
            ar += d.x * c + d.y * s;
            ai += d.y * c - d.x * s;
        }
        __syncthreads();
    }
    if (n < N) out[n] = make_double2(ar, ai);
}

}  // namespace

extern "C" int NDFT2_2D_GPU(void *handle, const double *X, const double *Y, const double *u, const double *v,
                            const nfft2d_c128 *in, nfft2d_c128 *out, int Nu, int Nv, int M, int N)
{
    (void)handle;
    if (M < 0 || N < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: negative size");
    if (Nu == 0 || Nv == 0) return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: Nu and Nv must be non-zero");
    if (N == 0) return NFFT2D_OK;
    if (!u || !v || !out || (M > 0 && (!X || !Y || !in)))
        return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: null pointer");
    // (u, v) a grid: FP64 MFMA contraction; otherwise the direct sum (each kernel
    // reads the detector's finding and returns at once when it is not its case)
    // point ranges: enough work items for the whole chip at the supertile count a
    // square grid of N outputs would have, each range at least 16 points
    const long long nsup_est = std::max<long long>(1, (long long)N / (kSupRows * kSupCols));
    int ranges = (int)std::min<long long>(kMaxRanges, std::max<long long>(1, kGridWavesTotal / nsup_est));
    ranges = std::max(1, std::min(ranges, (M + 15) / 16));
    const int chunk = ((M + ranges - 1) / ranges + 3) & ~3;   // a multiple of 4 points
    const int dev = [] { int d = 0; (void)hipGetDevice(&d); return d; }();
    double2 *part = nullptr;
    // held until the launches are queued: a call from another host thread that
    // grows the workspace frees the old one only after a device synchronise
    std::lock_guard<std::mutex> lk(g_ws_mu);
    {
        GridWs &ws = g_ws[dev];
        // partial sums, then (from an even double2 index: 32-byte aligned) the steps
        const size_t need = (((size_t)ranges * (size_t)N + 1) & ~(size_t)1) + 2 * (size_t)M;
        if (ws.count < need) {
            if (ws.ptr) {
                (void)hipDeviceSynchronize();
                (void)hipFree(ws.ptr);
                ws.ptr = nullptr;
                ws.count = 0;
            }
            if (nfft2d_malloc(&ws.ptr, sizeof(double2) * need) != hipSuccess) {
                (void)hipGetLastError();
                ws.ptr = nullptr;
                return nfft2d_fail(NFFT2D_ERR_ALLOC, "NDFT2_2D_GPU: workspace of %zu bytes", sizeof(double2) * need);
            }
            ws.count = need;
        }
        part = ws.ptr;
    }
    double4 *steps = reinterpret_cast<double4 *>(part + (((size_t)ranges * (size_t)N + 1) & ~(size_t)1));
    hipLaunchKernelGGL(k_grid_detect, dim3(1), dim3(kDetThreads), 0, nullptr, u, v, N);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grid_verify, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, nullptr, u, v, N);
    NFFT2D_LAUNCH_CHECK();
    if (M > 0) {
        hipLaunchKernelGGL(k_ndft_steps, dim3((M + kThreads - 1) / kThreads), dim3(kThreads), 0, nullptr, X, Y,
                           1.0 / Nu, 1.0 / Nv, M, steps);
        NFFT2D_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_ndft_grid, dim3(kGridWavesTotal / 4), dim3(256), 0, nullptr, X, Y,
                       reinterpret_cast<const double2 *>(in), steps, part, 1.0 / Nu, 1.0 / Nv, M, ranges, chunk);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ndft_grid_sum, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, nullptr, part, ranges,
                       reinterpret_cast<double2 *>(out), N);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ndft_direct, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, nullptr, X, Y, u, v,
                       reinterpret_cast<const double2 *>(in), reinterpret_cast<double2 *>(out), 1.0 / Nu, 1.0 / Nv, M,
                       N);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_grid_info(int *is_grid, int *nu, int *nv, int *vmajor)
{
    if (!is_grid || !nu || !nv || !vmajor) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_grid_info: null pointer");
    GridInfo g;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpyFromSymbol(&g, HIP_SYMBOL(g_grid), sizeof(g), 0, hipMemcpyDeviceToHost) != hipSuccess)
        return nfft2d_fail(NFFT2D_ERR_HIP, "nfft2d_ndft_grid_info: reading the detector's finding failed");
    *is_grid = g.ok;
    *nu = g.nu;
    *nv = g.nv;
    *vmajor = g.vmajor;
    return NFFT2D_OK;
}

// ============================================================================
// Part 2: the type-2 direct sum as a split-precision MFMA contraction
// (kernels in ndft2d_mfma.hpp).  A plan holds the mode operand.
namespace {

constexpr int kNdftMaxN = 8192;
constexpr int kNdftC = ndft::kChunkTiles;  // MFMA column tiles per chunk

size_t ndft_slice_bytes(int prec)
{
    switch (prec) {
    case NFFT2D_NDFT_FP32: return sizeof(ndft::Slice<ndft::kNdftFp32>);
    case NFFT2D_NDFT_BF16X3: return sizeof(ndft::Slice<ndft::kNdftBf16x3>);
    default: return sizeof(ndft::Slice<ndft::kNdftFp16>);
    }
}

}  // namespace

struct nfft2d_ndft_plan_s {
    int N1, N2, prec, device;
    int nch, nks;
    void *G;              // mode operand, nch * nks * C * 64 slices
    float *scale;         // N1 per-row scales and their inverses
    float *inv_scale;
};

extern "C" int nfft2d_ndft_plan_create(nfft2d_ndft_plan **plan, int N1, int N2, int precision)
{
    if (!plan) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_create: null plan pointer");
    *plan = nullptr;
    if (N1 <= 0 || N2 <= 0 || (N1 & 1) || (N2 & 1) || N1 > kNdftMaxN || N2 > kNdftMaxN)
        return nfft2d_fail(NFFT2D_ERR_SIZE, "nfft2d_ndft_plan_create: N1=%d N2=%d must be even, in [2, %d]", N1, N2,
                           kNdftMaxN);
    if (precision != NFFT2D_NDFT_FP32 && precision != NFFT2D_NDFT_BF16X3 && precision != NFFT2D_NDFT_FP16)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_create: unknown precision %d", precision);
    nfft2d_ndft_plan_s *p = new (std::nothrow) nfft2d_ndft_plan_s();
    if (!p) return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft2d_ndft_plan_create: host allocation failed");
    p->N1 = N1;
    p->N2 = N2;
    p->prec = precision;
    NFFT2D_HIP_TRY(hipGetDevice(&p->device));
    p->nch = (N1 + ndft::kChunkK1 - 1) / ndft::kChunkK1;
    p->nks = (N2 + 8 * ndft::kKsPad - 1) / (8 * ndft::kKsPad) * ndft::kKsPad;
    const size_t gbytes = (size_t)p->nch * p->nks * kNdftC * 64 * ndft_slice_bytes(precision);
    if (nfft2d_malloc(&p->G, gbytes) != hipSuccess || nfft2d_malloc(&p->scale, sizeof(float) * N1) != hipSuccess ||
        nfft2d_malloc(&p->inv_scale, sizeof(float) * N1) != hipSuccess) {
        (void)hipGetLastError();
        nfft2d_ndft_plan_destroy(p);
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft2d_ndft_plan_create: device allocation of %zu bytes failed", gbytes);
    }
    *plan = p;
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_plan_destroy(nfft2d_ndft_plan *p)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_destroy: null plan");
    if (p->G) (void)hipFree(p->G);
    if (p->scale) (void)hipFree(p->scale);
    if (p->inv_scale) (void)hipFree(p->inv_scale);
    delete p;
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_set_modes(nfft2d_ndft_plan *p, const nfft2d_c64 *data, void *stream)
{
    if (!p || !data) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_set_modes: null argument");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const float2 *F = reinterpret_cast<const float2 *>(data);
    hipLaunchKernelGGL(ndft::k_ndft_rowscale, dim3(p->N1), dim3(256), 0, s, F, p->scale, p->inv_scale, p->N2,
                       p->prec == NFFT2D_NDFT_FP16 ? 1 : 0);
    NFFT2D_LAUNCH_CHECK();
    const long long total = (long long)p->nch * p->nks * kNdftC * 64;
    const dim3 grid((unsigned)((total + 255) / 256));
    switch (p->prec) {
    case NFFT2D_NDFT_FP32:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftFp32, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftFp32> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    case NFFT2D_NDFT_BF16X3:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftBf16x3, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftBf16x3> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    default:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftFp16, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftFp16> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_eval(nfft2d_ndft_plan *p, nfft2d_c64 *result, const float *x, const float *y, int M,
                                void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: null plan");
    if (M < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: negative M");
    if (M == 0) return NFFT2D_OK;
    if (!result || !x || !y) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: null pointer");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((M + ndft::kPtsPerGroup - 1) / ndft::kPtsPerGroup));
    float2 *out = reinterpret_cast<float2 *>(result);
#ifndef NDFT_SPLIT
#define NDFT_SPLIT 2
#endif
    const dim3 block(64 * ndft::kWaves * NDFT_SPLIT);
    switch (p->prec) {
    case NFFT2D_NDFT_FP32:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftFp32, 4, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftFp32> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    case NFFT2D_NDFT_BF16X3:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftBf16x3, 16, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftBf16x3> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    default:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftFp16, 16, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftFp16> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_type2(nfft2d_ndft_plan *p, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x,
                                 const float *y, int M, void *stream)
{
    int rc = nfft2d_ndft_set_modes(p, data, stream);
    if (rc != NFFT2D_OK) return rc;
    return nfft2d_ndft_eval(p, result, x, y, M, stream);
}
