// nufft3.hip -- the type-3 2-D NUFFT (non-uniform to non-uniform, FP64) on
// gfx950:
//     F_j = sum_k f_k e^{-i (s_j x_k + t_j y_k)},
// the GPU counterpart of NUFFT_2D_CPU (NFFT/NFFT3/C/NFFT3_2D_C.cpp:199-312):
// Gaussian gridding with the reference's parameters (R = 2.2, the tap
// half-width msp from eps, an Mry x Mrx grid sized by the data extents).
//
// One call:
//   1. k3_extent / k3_extent_final: min and max of x, y, s, t (and whether all
//      are finite).  The host reads them back -- the call's one
//      synchronisation -- and derives every parameter exactly as :214-267.
//   2. k3_prep_in / k3_prep_out: per input its pre-modulated sample and grid
//      cell, per output its cell; both are counted into bins (inputs 8 x 8
//      cells, outputs 16 x 16 cells and row) and sorted (k3_scan_*, k3_scatter*).
//   3. k3_spread: one block per 16 x 16 tile of the grid sums
//      c_k ey_k(j) ex_k(i) over the inputs within reach of the tile, four
//      inputs per v_mfma_f64_16x16x4f64 (the Gaussian of :76 is separable);
//      the cell factor fact e^{b (Dx^2 dsi^2 + Dy^2 dsj^2)} multiplies the
//      finished tile.  Each cell is written once: no atomics, no zero fill.
//   4. The forward 2-D DFT of the grid (:280-292): a pass over the rows, a
//      transpose, a pass over the rows of the transpose.  Each line is one
//      LDS-resident radix-4 Stockham FFT; lengths that are not powers of two
//      go through Bluestein's chirp-z convolution at the next power of two
//      >= 2n - 1.
//   5. k3_interp: one block per output bin stages the bin's grid patch in LDS
//      with the circular and FFT shifts and the checkerboard sign of
//      circAndFFTShiftsFunction (:135-151) folded into the index; one wave
//      per output sums the (2 msp + 1)^2 separable Gaussian taps and applies
//      the post-compensation of spectralConvolutionFunctionCompensation
//      (:156-194).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>

#include "../../include/nfft2d.h"
#include "nfft2d_internal.hpp"

namespace nufft3 {

constexpr int kTile = 16;              // grid tile and output bin edge, cells
constexpr int kInBin = 8;              // input bin edge, cells (finer: fewer out-of-reach candidates per tile)
constexpr int kMaxMsp = 42;            // LDS patch (16 + 2 msp)^2 x 16 B <= 160 000 B
constexpr int kMaxFft = 8192;          // longest LDS-resident FFT (128 KB)
constexpr int kExtentBlocks = 1024;
typedef double d4 __attribute__((ext_vector_type(4)));

struct Par {
    double xb, yb, sb, tb, kkx, kky, Dx, Dy, b, t1, fact;
    int msp, Mrx, Mry, nbx, nby;   // grid, and its 16 x 16-cell tiles (spread) / output bins (interpolation)
    int ibx, iby;                   // 8 x 8-cell input bins
};

struct alignas(16) InRec {     // one input, in bin order
    double xt, yt, re, im;     // scaled offsets (x - xb) / Dx, (y - yb) / Dy; pre-modulated sample
    int nno, mmo, pad0, pad1;  // its cell
};
struct alignas(16) OutRec {    // one output, in bin order
    double difs, dift, sd, td; // fractional cell offsets; s - sb, t - tb (rescaled)
    int nsn, ntn, idx, pad;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ------------------------------------------------------------ 1. extents
// 9 values per block: max x, min x, max y, min y, max s, min s, max t, min t,
// and 1.0 if any coordinate is not finite.  The start values are -FLT_MAX /
// FLT_MAX, as the reference's thrust::reduce calls (:214-231).
__device__ __forceinline__ double wmax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wmin(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(256) void k3_extent(const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ s, const double *__restrict__ t,
                                                 int len_in, int len_out, double *__restrict__ part)
{
    __shared__ double red[4][9];
    double v[9];
    for (int q = 0; q < 8; q += 2) {
        v[q] = -(double)FLT_MAX;
        v[q + 1] = (double)FLT_MAX;
    }
    v[8] = 0.0;
    const int stride = gridDim.x * 256;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < len_in; i += stride) {
        const double a = x[i], c = y[i];
        if (!isfinite(a) || !isfinite(c)) v[8] = 1.0;
        if (a > v[0]) v[0] = a;
        if (a < v[1]) v[1] = a;
        if (c > v[2]) v[2] = c;
        if (c < v[3]) v[3] = c;
    }
    for (int i = blockIdx.x * 256 + threadIdx.x; i < len_out; i += stride) {
        const double a = s[i], c = t[i];
        if (!isfinite(a) || !isfinite(c)) v[8] = 1.0;
        if (a > v[4]) v[4] = a;
        if (a < v[5]) v[5] = a;
        if (c > v[6]) v[6] = c;
        if (c < v[7]) v[7] = c;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double r = (q & 1) && q < 8 ? wmin(v[q]) : wmax(v[q]);
        if (lane == 0) red[w][q] = r;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int q = threadIdx.x;
        double r = red[0][q];
        for (int ww = 1; ww < 4; ++ww) r = (q & 1) && q < 8 ? fmin(r, red[ww][q]) : fmax(r, red[ww][q]);
        part[blockIdx.x * 9 + q] = r;
    }
}

__global__ __launch_bounds__(64) void k3_extent_final(const double *__restrict__ part, int nblk,
                                                      double *__restrict__ out)
{
    for (int q = 0; q < 9; ++q) {
        const bool mn = (q & 1) && q < 8;
        double r = mn ? (double)FLT_MAX : -(double)FLT_MAX;
        if (q == 8) r = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 64) r = mn ? fmin(r, part[b * 9 + q]) : fmax(r, part[b * 9 + q]);
        r = mn ? wmin(r) : wmax(r);
        if (threadIdx.x == 0) out[q] = r;
    }
}

// ------------------------------------------------------------ 2. points and bins
__device__ __forceinline__ int bin_of(int i, int j, const Par &p)
{
    return (j / kTile) * p.nbx + i / kTile;
}

__global__ __launch_bounds__(256) void k3_prep_in(const double *__restrict__ x0, const double *__restrict__ y0,
                                                  const double2 *__restrict__ f, int n, Par p,
                                                  InRec *__restrict__ rec, int2 *__restrict__ kr,
                                                  int *__restrict__ cnt)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    // spatialConvolutionFunction (:38-63), on the rescaled x / kk_x, y / kk_y (:247-248)
    const double x = x0[k] / p.kkx, y = y0[k] / p.kky;
    const double arg = p.sb * x + p.tb * y;
    double sn, cs;
    sincos(arg, &sn, &cs);
    const double2 v = f[k];
    InRec r;
    r.re = v.x * cs + v.y * sn;
    r.im = v.y * cs - v.x * sn;
    r.xt = (x - p.xb) / p.Dx;
    r.yt = (y - p.yb) / p.Dy;
    r.nno = (int)floor((double)(p.Mrx / 2) + r.xt);
    r.mmo = (int)floor((double)(p.Mry / 2) + r.yt);
    r.pad0 = r.pad1 = 0;
    // the extents put every cell in [Mrx / 4 - 1, 3 Mrx / 4]; the bound keeps
    // a rounding at the edge from leaving the grid
    const int nb = p.ibx * p.iby;
    const bool ok = r.nno >= 0 && r.nno < p.Mrx && r.mmo >= 0 && r.mmo < p.Mry;
    const int bin = ok ? (r.mmo / kInBin) * p.ibx + r.nno / kInBin : nb;   // bin nb is never read
    rec[k] = r;
    kr[k] = make_int2(bin, atomicAdd(&cnt[bin], 1));
}

__global__ __launch_bounds__(256) void k3_prep_out(const double *__restrict__ s0, const double *__restrict__ t0,
                                                   int n, Par p, OutRec *__restrict__ rec,
                                                   int2 *__restrict__ kr, int *__restrict__ cnt,
                                                   double2 *__restrict__ F)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    // spectralConvolutionFunctionCompensation (:164-167), s * kk_x, t * kk_y (:249-250), Ds = Dt = 1
    const double s = s0[k] * p.kkx, t = t0[k] * p.kky;
    OutRec r;
    r.sd = s - p.sb;
    r.td = t - p.tb;
    r.nsn = (int)floor((double)(p.Mrx / 2) + r.sd);
    r.difs = ((double)(p.Mrx / 2) + r.sd) - r.nsn;
    r.ntn = (int)floor((double)(p.Mry / 2) + r.td);
    r.dift = ((double)(p.Mry / 2) + r.td) - r.ntn;
    r.idx = k;
    r.pad = 0;
    const int nb = p.nbx * p.nby * kTile;   // keys: bin and the row inside it
    const bool ok = r.nsn >= 0 && r.nsn < p.Mrx && r.ntn >= 0 && r.ntn < p.Mry;
    const int bin = ok ? bin_of(r.nsn, r.ntn, p) * kTile + r.ntn % kTile : nb;
    if (!ok) F[k] = make_double2(NAN, NAN);
    rec[k] = r;
    kr[k] = make_int2(bin, atomicAdd(&cnt[bin], 1));
}

// Exclusive scan of the key counts in three steps: k3_scan_local scans each
// 1024-key segment in LDS and writes the segment totals, k3_scan_totals scans
// those (one block, a run of totals per thread), k3_scan_add adds them back.
// start[n] is the total.
__global__ __launch_bounds__(1024) void k3_scan_local(const int *__restrict__ cnt, int n, int *__restrict__ start,
                                                      int *__restrict__ tot)
{
    __shared__ int sums[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    const int v = i < n ? cnt[i] : 0;
    sums[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int u = threadIdx.x >= o ? sums[threadIdx.x - o] : 0;
        __syncthreads();
        sums[threadIdx.x] += u;
        __syncthreads();
    }
    if (i < n) start[i] = sums[threadIdx.x] - v;
    if (threadIdx.x == 1023) tot[blockIdx.x] = sums[1023];
}

__global__ __launch_bounds__(1024) void k3_scan_totals(int *__restrict__ tot, int n)
{
    __shared__ int sums[1024];
    const int per = (n + 1023) / 1024;
    const int lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
    int acc = 0;
    for (int i = lo; i < hi; ++i) acc += tot[i];
    sums[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int u = threadIdx.x >= o ? sums[threadIdx.x - o] : 0;
        __syncthreads();
        sums[threadIdx.x] += u;
        __syncthreads();
    }
    int run = sums[threadIdx.x] - acc;
    for (int i = lo; i < hi; ++i) {
        const int v = tot[i];
        tot[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) tot[n] = sums[1023];
}

__global__ __launch_bounds__(256) void k3_scan_add(int *__restrict__ start, int n, const int *__restrict__ tot)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) start[i] += tot[i >> 10];
    if (i == n) start[n] = tot[(n + 1023) >> 10];
}

template <typename R>
__global__ __launch_bounds__(256) void k3_scatter(const R *__restrict__ rec, const int2 *__restrict__ kr, int n,
                                                  const int *__restrict__ st, R *__restrict__ out)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int2 e = kr[k];
    out[st[e.x] + e.y] = rec[k];
}

// The inputs' scatter also writes their cells, in the same order, to an
// 8-byte array: the spread's reach test reads only that.
__global__ __launch_bounds__(256) void k3_scatter_in(const InRec *__restrict__ rec, const int2 *__restrict__ kr,
                                                     int n, const int *__restrict__ st, InRec *__restrict__ out,
                                                     int2 *__restrict__ cells)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int2 e = kr[k];
    const InRec r = rec[k];
    const int pos = st[e.x] + e.y;
    out[pos] = r;
    cells[pos] = make_int2(r.nno, r.mmo);
}

// ------------------------------------------------------------ 3. Gaussian gridding
// Block per 16 x 16 tile (tx, ty) of the row-major Mry x Mrx grid: cells
// i in [16 tx, 16 tx + 16), j in [16 ty, 16 ty + 16).  An input k reaches cell
// (i, j) when |i - nno_k| <= msp and |j - mmo_k| <= msp (:65-72).  The 4 waves
// take turns over the 64-input chunks of the 8 x 8-cell input bins within
// ceil(msp / 8) of the tile (the reach test reads the 8-byte cell array; only
// inputs within reach load their record) and sum their partial tiles in LDS
// at the end, in wave order.
// In a chunk, each input that reaches the tile writes its 16 row and 16
// column Gaussian factors for the tile to a slot in LDS (zero outside its
// reach), from two exponentials per dimension and a product recurrence (fast
// Gaussian gridding, Greengard & Lee, SIAM Rev. 46 (2004)): with
// u = yt - (j0 - Mry/2),
//     e^{-t1 (u - r - 1)^2} = e^{-t1 (u - r)^2} rho_r,
//     rho_r = e^{2 t1 (u - r) - t1},  rho_{r+1} = rho_r e^{-2 t1}.
// The slots are then taken four per MFMA step: lane l supplies
// A[r][k] = c_k ey_k(j0 + r) and B[k][r] = ex_k(i0 + r) at r = l % 16,
// k = l / 16, and accumulator element q of lane l is cell
// (j0 + 4 q + l / 16, i0 + l % 16).
constexpr int kSpreadWaves = 4;
constexpr int kChunk = 64;        // candidate inputs per wave and round
constexpr int kRow = kTile + 1;   // padded slot row (doubles): lane-strided writes spread over the banks

__global__ __launch_bounds__(64 * kSpreadWaves) void k3_spread(const InRec *__restrict__ pts,
                                                                const int2 *__restrict__ cells,
                                                                const int *__restrict__ st, Par p,
                                                                double2 *__restrict__ G)
{
    __shared__ union {
        struct {
            double ey[kSpreadWaves][kChunk][kRow], ex[kSpreadWaves][kChunk][kRow];
            double2 c[kSpreadWaves][kChunk];
        } s;
        double2 part[kSpreadWaves][kTile * kTile];
    } sh;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tx = blockIdx.x % p.nbx, ty = blockIdx.x / p.nbx;
    const int i0 = tx * kTile, j0 = ty * kTile;
    const int r = lane & 15, ks = lane >> 4;
    const double oy = (double)(j0 - p.Mry / 2), ox = (double)(i0 - p.Mrx / 2);
    const double kap = exp(-2.0 * p.t1);
    const int msp = p.msp, reach = (msp + kInBin - 1) / kInBin;   // in input bins
    const int rows = min(kTile, p.Mry - j0), cols = min(kTile, p.Mrx - i0);   // cells inside the grid
    d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
    constexpr int kSub = kTile / kInBin;   // input bins per tile side
    const int by0 = max(0, kSub * ty - reach), by1 = min(p.iby - 1, kSub * ty + kSub - 1 + reach);
    const int bx0 = max(0, kSub * tx - reach), bx1 = min(p.ibx - 1, kSub * tx + kSub - 1 + reach);
    const int2 far = make_int2(-(1 << 29), -(1 << 29));   // reaches nothing
    for (int by = by0; by <= by1; ++by) {
        const int s0 = st[by * p.ibx + bx0], s1 = st[by * p.ibx + bx1 + 1];   // the row's bins are contiguous
        int base = s0 + kChunk * w;
        int2 nxt = base + lane < s1 ? cells[base + lane] : far;
        for (; base < s1; base += kChunk * kSpreadWaves) {
            const int2 cell = nxt;
            const int after = base + kChunk * kSpreadWaves + lane;
            nxt = after < s1 ? cells[after] : far;
            const bool in = cell.x + msp >= i0 && cell.x - msp < i0 + kTile && cell.y + msp >= j0 &&
                            cell.y - msp < j0 + kTile;
            const unsigned long long mask = __ballot(in);
            if (mask == 0ull) continue;
            if (in) {
                const InRec *ep = pts + base + lane;   // only inputs within reach load their record
                const double2 xy = *reinterpret_cast<const double2 *>(&ep->xt);
                const double2 cv = *reinterpret_cast<const double2 *>(&ep->re);
                const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                sh.s.c[w][pos] = cv;
                // rows: reach [mmo - msp, mmo + msp] relative to j0, cells inside the grid
                const int ylo = max(0, cell.y - msp - j0), yhi = min(rows - 1, cell.y + msp - j0);
                const int xlo = max(0, cell.x - msp - i0), xhi = min(cols - 1, cell.x + msp - i0);
                const double uy = xy.y - oy, ux = xy.x - ox;
                double gy = exp(-p.t1 * (uy * uy)), ry = exp(2.0 * p.t1 * uy - p.t1);
                double gx = exp(-p.t1 * (ux * ux)), rx = exp(2.0 * p.t1 * ux - p.t1);
                double *py = sh.s.ey[w][pos], *px = sh.s.ex[w][pos];
#pragma unroll
                for (int q = 0; q < kTile; ++q) {
                    py[q] = q >= ylo && q <= yhi ? gy : 0.0;
                    px[q] = q >= xlo && q <= xhi ? gx : 0.0;
                    gy *= ry;
                    ry *= kap;
                    gx *= rx;
                    rx *= kap;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int n = __popcll(mask);
            for (int q = 0; q < n; q += 4) {
                const int sl = q + ks;
                double ar = 0.0, ai = 0.0, bb = 0.0;
                if (sl < n) {
                    const double2 c = sh.s.c[w][sl];
                    const double wy = sh.s.ey[w][sl][r];
                    ar = c.x * wy;
                    ai = c.y * wy;
                    bb = sh.s.ex[w][sl][r];
                }
                accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bb, accr, 0, 0, 0);
                acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bb, acci, 0, 0, 0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();   // slot reads done before the next chunk's writes
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();   // every wave's slot reads are done: the partial tiles may overwrite them
#pragma unroll
    for (int q = 0; q < 4; ++q) sh.part[w][(4 * q + ks) * kTile + r] = make_double2(accr[q], acci[q]);
    __syncthreads();
    if (threadIdx.x < kTile * kTile) {
        const int cell = threadIdx.x, ii = i0 + cell % kTile, jj = j0 + cell / kTile;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int ww = 0; ww < kSpreadWaves; ++ww) {   // fixed order
            const double2 v = sh.part[ww][cell];
            sr += v.x;
            si += v.y;
        }
        if (jj < p.Mry && ii < p.Mrx) {
            const double ddi = (double)(ii - p.Mrx / 2), ddj = (double)(jj - p.Mry / 2);
            const double sc = p.fact * exp(p.b * (p.Dy * p.Dy * ddj * ddj) + p.b * (p.Dx * p.Dx * ddi * ddi));
            G[(size_t)jj * p.Mrx + ii] = make_double2(sr * sc, si * sc);
        }
    }
}

// ------------------------------------------------------------ 4. FFTs of the grid lines
// In-place radix-4 (and one radix-2) Stockham FFT of m = 2^a points in LDS,
// forward (e^{-2 pi i jk / m}), after Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors" (SC 2008): stage with
// span Ns reads v_r = X[j + r m/R], twiddles it by e^{-2 pi i r (j mod Ns) /
// (Ns R)}, and writes the R-point DFT of v to X[(j - j mod Ns) R + j mod Ns
// + r Ns].  All reads of a stage precede its writes (barrier), so one buffer
// serves.  Twiddles e^{-2 pi i t / m} = T1[t / 64] T2[t % 64] from two small
// LDS tables.  Every thread holds at most 8 values.
struct TwLds {
    const double2 *t1, *t2;
    __device__ double2 operator[](int t) const { return cmul(t1[t >> 6], t2[t & 63]); }
};

__device__ void lds_fft(double2 *X, int m, const TwLds tw)
{
    const int T = blockDim.x, tid = threadIdx.x;
    for (int Ns = 1; Ns < m;) {
        const int R = (m / Ns) >= 4 ? 4 : 2;
        const int nb = m / R, per = (nb + T - 1) / T;   // butterflies per thread: <= 2 (radix 4), <= 4 (radix 2)
        const int tstep = m / (Ns * R);
        double2 v[8];
        if (R == 4) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = tid + u * T;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1);
                    double2 a0 = X[j], a1 = X[j + nb], a2 = X[j + 2 * nb], a3 = X[j + 3 * nb];
                    if (Ns > 1) {
                        a1 = cmul(a1, tw[k * tstep]);
                        a2 = cmul(a2, tw[2 * k * tstep]);
                        a3 = cmul(a3, tw[3 * k * tstep]);
                    }
                    const double2 b0 = make_double2(a0.x + a2.x, a0.y + a2.y);
                    const double2 b1 = make_double2(a0.x - a2.x, a0.y - a2.y);
                    const double2 b2 = make_double2(a1.x + a3.x, a1.y + a3.y);
                    const double2 b3 = make_double2(a1.y - a3.y, a3.x - a1.x);   // (a1 - a3) (-i)
                    v[4 * u + 0] = make_double2(b0.x + b2.x, b0.y + b2.y);
                    v[4 * u + 1] = make_double2(b1.x + b3.x, b1.y + b3.y);
                    v[4 * u + 2] = make_double2(b0.x - b2.x, b0.y - b2.y);
                    v[4 * u + 3] = make_double2(b1.x - b3.x, b1.y - b3.y);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = tid + u * T;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1), o = (j - k) * 4 + k;
                    X[o] = v[4 * u];
                    X[o + Ns] = v[4 * u + 1];
                    X[o + 2 * Ns] = v[4 * u + 2];
                    X[o + 3 * Ns] = v[4 * u + 3];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = tid + u * T;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1);
                    double2 a0 = X[j], a1 = X[j + nb];
                    if (Ns > 1) a1 = cmul(a1, tw[k * tstep]);
                    v[2 * u] = make_double2(a0.x + a1.x, a0.y + a1.y);
                    v[2 * u + 1] = make_double2(a0.x - a1.x, a0.y - a1.y);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = tid + u * T;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1), o = (j - k) * 2 + k;
                    X[o] = v[2 * u];
                    X[o + Ns] = v[2 * u + 1];
                }
            }
        }
        __syncthreads();
        Ns *= R;
    }
}

// Forward DFT of `lines` contiguous lines of n points (line stride n), one
// line per block, m = n (n a power of two) or Bluestein at m >= 2n - 1:
//   X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),  w_j = e^{-i pi j^2 / n},
// the convolution taken as IFFT(FFT(x w) * bh), bh = FFT of the wrapped
// conj(w) sequence, and IFFT(Y) = conj(FFT(conj(Y))) / m.
// LDS: m data values, then the twiddle tables T1 (m / 64 entries, at least
// 1) and T2 (64).
__host__ __device__ constexpr int fft_lds_entries(int m) { return m + (m >= 64 ? m / 64 : 1) + 64; }

__global__ __launch_bounds__(1024) void k3_fft_lines(double2 *__restrict__ data, int n, int m,
                                                     const double2 *__restrict__ twg,
                                                     const double2 *__restrict__ chirp,
                                                     const double2 *__restrict__ bh)
{
    extern __shared__ double2 X[];
    double2 *t1 = X + m, *t2 = t1 + (m >= 64 ? m / 64 : 1);
    double2 *row = data + (size_t)blockIdx.x * n;
    const int T = blockDim.x;
    for (int k = threadIdx.x; k < 64; k += T) t2[k] = k < m ? twg[k] : make_double2(1.0, 0.0);
    for (int k = threadIdx.x; k < (m >= 64 ? m / 64 : 1); k += T) t1[k] = twg[64 * k];
    const TwLds tw{t1, t2};
    if (!chirp) {
        for (int k = threadIdx.x; k < m; k += T) X[k] = row[k];
        __syncthreads();
        lds_fft(X, m, tw);
        for (int k = threadIdx.x; k < m; k += T) row[k] = X[k];
        return;
    }
    for (int k = threadIdx.x; k < m; k += T) X[k] = k < n ? cmul(row[k], chirp[k]) : make_double2(0.0, 0.0);
    __syncthreads();
    lds_fft(X, m, tw);
    for (int k = threadIdx.x; k < m; k += T) {
        const double2 z = cmul(X[k], bh[k]);
        X[k] = make_double2(z.x, -z.y);
    }
    __syncthreads();
    lds_fft(X, m, tw);
    const double inv = 1.0 / m;
    for (int k = threadIdx.x; k < n; k += T) {
        const double2 z = X[k];
        row[k] = cmul(make_double2(z.x * inv, -z.y * inv), chirp[k]);
    }
}

// tw[t] = e^{-2 pi i t / m}, t < m.
__global__ __launch_bounds__(256) void k3_twiddles(int m, double2 *__restrict__ tw)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    double sn, cs;
    sincospi(2.0 * t / m, &sn, &cs);
    tw[t] = make_double2(cs, -sn);
}

// Bluestein tables of a length-n line: chirp[j] = e^{-i pi j^2 / n} (j < n;
// j^2 reduced mod 2n exactly) and b = the length-m sequence holding
// conj(chirp[j]) at j and at m - j, zero between (transformed in place
// afterwards).
__global__ __launch_bounds__(256) void k3_chirp(int n, int m, double2 *__restrict__ chirp, double2 *__restrict__ b)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    double2 bj = make_double2(0.0, 0.0);
    const int jj = j < n ? j : (j > m - n ? m - j : -1);
    if (jj >= 0) {
        const long long q = ((long long)jj * jj) % (2LL * n);
        double sn, cs;
        sincospi((double)q / n, &sn, &cs);
        if (j < n) chirp[j] = make_double2(cs, -sn);
        bj = make_double2(cs, sn);
    }
    b[j] = bj;
}

// out[c][r] = in[r][c], in: rows x cols.
__global__ __launch_bounds__(256) void k3_transpose(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                    int rows, int cols)
{
    __shared__ double2 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int rr = r0 + k, cc = c0 + tx;
        if (rr < rows && cc < cols) tile[k][tx] = in[(size_t)rr * cols + cc];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int cc = c0 + k, rr = r0 + tx;
        if (rr < rows && cc < cols) out[(size_t)cc * rows + rr] = tile[tx][k];
    }
}

// ------------------------------------------------------------ 5. interpolation at the outputs
// Block per output bin (by, bx), 8 waves.  The patch of the shifted spectrum
//     gs[mm][nn] = (-1)^(i + j) FT[j][i],  i = (nn + Mrx/2) mod Mrx,
//     j = (mm + Mry/2) mod Mry
// (FT the 2-D DFT, held transposed as FTt[i][j]) for mm in
// [16 by - msp, 16 by + 16 + msp) and nn likewise goes to LDS as P[c][a]
// (c = nn offset, a = mm offset, W = 16 + 2 msp per side; wrapped indices
// outside the grid are loaded but never weighted).  The outputs of the bin
// come sorted by their row ntn.  A wave takes 16 of them at a time and
// contracts, for each output j, H_j[a] = sum_c P[c][a] ex_j[c] on
// v_mfma_f64_16x16x4f64 (A[a][k] = P[c0 + k][a0 + a] from LDS,
// B[k][j] = ex_j[c0 + k], accumulator element q of lane l = H_{l % 16}[a0 + 4 q
// + l / 16]) over the rows and columns that the 16 windows cover, then
// F_j = sum_a ey_j[a] H_j[a] and the post-compensation.  Along c the taps
// follow e^{-t1 (d - c - 4)^2} = e^{-t1 (d - c)^2} rho_c, rho_{c+4} = rho_c
// e^{-32 t1} (two multiplications per step).
constexpr int kInterpWaves = 8;

__global__ __launch_bounds__(64 * kInterpWaves) void k3_interp(const double2 *__restrict__ FTt,
                                                                const OutRec *__restrict__ outs,
                                                                const int *__restrict__ st, Par p,
                                                                double2 *__restrict__ F)
{
    extern __shared__ double2 P[];
    const int bin = blockIdx.x;
    const int s0 = st[bin * kTile], s1 = st[bin * kTile + kTile];   // the bin's 16 row keys
    if (s0 == s1) return;
    const int bx = bin % p.nbx, by = bin / p.nbx;
    const int msp = p.msp, W = kTile + 2 * msp, taps = 2 * msp + 1;
    const int mm0 = by * kTile - msp, nn0 = bx * kTile - msp;
    for (int e = threadIdx.x; e < W * W; e += 64 * kInterpWaves) {
        const int a = e % W, c = e / W;   // a: mm offset (contiguous in FTt), c: nn offset
        int mm = mm0 + a, nn = nn0 + c;
        mm += mm < 0 ? p.Mry : (mm >= p.Mry ? -p.Mry : 0);
        nn += nn < 0 ? p.Mrx : (nn >= p.Mrx ? -p.Mrx : 0);
        int i = nn + p.Mrx / 2, j = mm + p.Mry / 2;
        i -= i >= p.Mrx ? p.Mrx : 0;
        j -= j >= p.Mry ? p.Mry : 0;
        double2 v = FTt[(size_t)i * p.Mry + j];
        if ((i + j) & 1) v = make_double2(-v.x, -v.y);
        P[c * W + a] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int jo = lane & 15, kq = lane >> 4;
    const double kappa = exp(-32.0 * p.t1);
    for (int g = s0 + kTile * w; g < s1; g += kTile * kInterpWaves) {
        const int o = g + jo;
        const bool valid = o < s1;
        const OutRec rec = outs[valid ? o : s1 - 1];
        const int ra = rec.ntn - msp - mm0, ca = rec.nsn - msp - nn0;   // window origin in the patch, in [0, 16)
        int rlo = valid ? ra : 1 << 20, rhi = valid ? ra : -(1 << 20);
        int clo = valid ? ca : 1 << 20, chi = valid ? ca : -(1 << 20);
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            rlo = min(rlo, __shfl_xor(rlo, off, 64));
            rhi = max(rhi, __shfl_xor(rhi, off, 64));
            clo = min(clo, __shfl_xor(clo, off, 64));
            chi = max(chi, __shfl_xor(chi, off, 64));
        }
        rhi += taps - 1;
        chi += taps - 1;
        const double dxc = rec.difs + ca + msp;   // ex_j[c] = e^{-t1 (dxc - c)^2} on c in [ca, ca + 2 msp]
        const double dya = rec.dift + ra + msp;   // ey_j[a] = e^{-t1 (dya - a)^2} on a in [ra, ra + 2 msp]
        double sr = 0.0, si = 0.0;
        for (int a0 = rlo; a0 <= rhi; a0 += 16) {
            d4 hr = {0.0, 0.0, 0.0, 0.0}, hi = {0.0, 0.0, 0.0, 0.0};
            const int arow = a0 + jo;
            double dc = dxc - (double)(clo + kq);
            double ev = exp(-p.t1 * (dc * dc)), rho = exp(8.0 * p.t1 * dc - 16.0 * p.t1);
            for (int c0 = clo; c0 <= chi; c0 += 4) {
                const int c = c0 + kq;
                const double bv = valid && c >= ca && c < ca + taps ? ev : 0.0;
                const double2 av = arow < W && c < W ? P[c * W + arow] : make_double2(0.0, 0.0);
                hr = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv, hr, 0, 0, 0);
                hi = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv, hi, 0, 0, 0);
                ev *= rho;
                rho *= kappa;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int a = a0 + 4 * q + kq;
                if (a >= ra && a < ra + taps) {
                    const double da = dya - (double)a;
                    const double ey = exp(-p.t1 * (da * da));
                    sr = fma(ey, hr[q], sr);
                    si = fma(ey, hi[q], si);
                }
            }
        }
        sr += __shfl_xor(sr, 16, 64);
        si += __shfl_xor(si, 16, 64);
        sr += __shfl_xor(sr, 32, 64);
        si += __shfl_xor(si, 32, 64);
        if (kq == 0 && valid) {
            const double a1 = p.Dy * rec.td, a2 = p.Dx * rec.sd;
            const double temp = exp(p.b * (a1 * a1 + a2 * a2));
            const double arg = rec.sd * p.xb + rec.td * p.yb;
            double sn, cs;
            sincos(arg, &sn, &cs);
            F[rec.idx] = make_double2((sr * cs + si * sn) * temp, (si * cs - sr * sn) * temp);
        }
    }
}

}  // namespace nufft3

using namespace nufft3;

// ============================================================== host side
namespace {

constexpr double kPi3 = 3.141592653589793238462643383279502884;

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int next_pow2(int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}
int fft_len(int n) { return is_pow2(n) ? n : next_pow2(2 * n - 1); }

// NUFFT_2D_CPU's parameters from the extents (:214-267), in its order of
// operations.  ext: max/min of x, y, s, t.
#pragma clang fp contract(off)
void derive(const double *ext, double eps, double *out)
{
    double Mx = ext[0], mx = ext[1], My = ext[2], my = ext[3];
    double Ms = ext[4], ms = ext[5], Mt = ext[6], mt = ext[7];
    double xb = (Mx + mx) / 2, X1 = fabs(Mx - xb);
    double yb = (My + my) / 2, Y1 = fabs(My - yb);
    double sb = (Ms + ms) / 2, S = fabs(Ms - sb);
    double tb = (Mt + mt) / 2, T = fabs(Mt - tb);
    const double R = 2.2;
    double msp = 2 + 2 * (R * R) * (-log(eps / 76)) / (kPi3 * (R * R - 2));
    const double mrx = 2 * ceil((X1 * S / kPi3) * R * R + R * msp);
    const double mry = 2 * ceil((Y1 * T / kPi3) * R * R + R * msp);
    double kkx = X1 / (kPi3 / 2), kky = Y1 / (kPi3 / 2);
    Mx /= kkx; mx /= kkx; xb = (Mx + mx) / 2;
    Ms *= kkx; ms *= kkx; sb = (Ms + ms) / 2;
    My /= kky; my /= kky; yb = (My + my) / 2;
    Mt *= kky; mt *= kky; tb = (Mt + mt) / 2;
    const double b = (msp - 2) / (4 * kPi3), t1 = 1 / (4 * b);
    msp = floor(msp - 2);
    double fact = 1. / (4. * kPi3 * b);
    out[0] = xb; out[1] = yb; out[2] = sb; out[3] = tb; out[4] = kkx; out[5] = kky;
    out[6] = 2 * kPi3 / mrx; out[7] = 2 * kPi3 / mry; out[8] = b; out[9] = t1; out[10] = fact * fact;
    out[11] = msp; out[12] = mrx; out[13] = mry;
}
#pragma clang fp contract(on)

struct Workspace {
    unsigned char *ptr = nullptr;
    size_t size = 0;
};
std::mutex g_mu;
std::map<int, Workspace> g_ws;

int ensure(Workspace &ws, size_t bytes)
{
    if (ws.size >= bytes) return NFFT2D_OK;
    if (ws.ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ws.ptr);
        ws.ptr = nullptr;
        ws.size = 0;
    }
    const size_t want = bytes + bytes / 4;
    if (nfft2d_malloc(&ws.ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        ws.ptr = nullptr;
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "type-3 workspace of %zu bytes", want);
    }
    ws.size = want;
    return NFFT2D_OK;
}

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

struct Carve {
    unsigned char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t n)
    {
        T *p = reinterpret_cast<T *>(base + off);
        off += al(n * sizeof(T));
        return p;
    }
};

int check_args(const void *x, const void *y, const void *s, const void *t, const void *f, const void *F,
               double eps, int len_in, int len_out)
{
    if (len_in < 0 || len_out < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "len_in = %d, len_out = %d", len_in, len_out);
    if (len_in > kMaxPoints || len_out > kMaxPoints)
        return nfft2d_fail(NFFT2D_ERR_ARG, "at most %d inputs and outputs", kMaxPoints);
    if (!(eps > 0.0 && eps < 1.0)) return nfft2d_fail(NFFT2D_ERR_ARG, "eps = %g: 0 < eps < 1 required", eps);
    if (len_out > 0 && (!s || !t || !F)) return nfft2d_fail(NFFT2D_ERR_ARG, "null output-side pointer");
    if (len_in > 0 && (!x || !y || !f)) return nfft2d_fail(NFFT2D_ERR_ARG, "null input-side pointer");
    return NFFT2D_OK;
}

// Extents on the device, parameters on the host; out[14] as orc_nufft3_params.
int params(Workspace &ws, const double *x, const double *y, const double *s, const double *t, double eps,
           int len_in, int len_out, double *out)
{
    const int L = std::max(len_in, len_out);
    const int nblk = std::max(1, std::min(kExtentBlocks, (L + 255) / 256));
    int rc = ensure(ws, al(sizeof(double) * 9 * (kExtentBlocks + 1)));
    if (rc) return rc;
    double *part = reinterpret_cast<double *>(ws.ptr), *fin = part + 9 * kExtentBlocks;
    hipLaunchKernelGGL(k3_extent, dim3(nblk), dim3(256), 0, nullptr, x, y, s, t, len_in, len_out, part);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k3_extent_final, dim3(1), dim3(64), 0, nullptr, part, nblk, fin);
    NFFT2D_LAUNCH_CHECK();
    double ext[9];
    NFFT2D_HIP_TRY(hipMemcpy(ext, fin, sizeof(ext), hipMemcpyDeviceToHost));
    if (ext[8] != 0.0) return nfft2d_fail(NFFT2D_ERR_ARG, "a coordinate is not finite");
    if (!(ext[0] > ext[1]) || !(ext[2] > ext[3]))
        return nfft2d_fail(NFFT2D_ERR_ARG,
                           "the inputs span no extent in x or y (all equal): the type-3 scaling divides by it");
    derive(ext, eps, out);
    return NFFT2D_OK;
}

int run(const double *x, const double *y, const double *s, const double *t, const nfft2d_c128 *f,
        nfft2d_c128 *F, double eps, int len_in, int len_out)
{
    int rc = check_args(x, y, s, t, f, F, eps, len_in, len_out);
    if (rc) return rc;
    if (len_out == 0) return NFFT2D_OK;
    if (len_in == 0) {   // the empty sum
        NFFT2D_HIP_TRY(hipMemsetAsync(F, 0, sizeof(nfft2d_c128) * (size_t)len_out, nullptr));
        return NFFT2D_OK;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    NFFT2D_HIP_TRY(hipGetDevice(&dev));
    Workspace &ws = g_ws[dev];
    double pv[14];
    if ((rc = params(ws, x, y, s, t, eps, len_in, len_out, pv))) return rc;
    if (!(pv[12] <= kMaxFft && pv[13] <= kMaxFft))
        return nfft2d_fail(NFFT2D_ERR_SIZE,
                           "grid %.0f x %.0f (Mrx x Mry): at most 4096 (8192 for powers of two) per side; reduce the "
                           "products of the extents X1 S and Y1 T",
                           pv[12], pv[13]);
    Par p;
    p.xb = pv[0]; p.yb = pv[1]; p.sb = pv[2]; p.tb = pv[3]; p.kkx = pv[4]; p.kky = pv[5];
    p.Dx = pv[6]; p.Dy = pv[7]; p.b = pv[8]; p.t1 = pv[9]; p.fact = pv[10];
    p.msp = (int)pv[11]; p.Mrx = (int)pv[12]; p.Mry = (int)pv[13];
    if (p.msp < 0 || p.msp > kMaxMsp)
        return nfft2d_fail(NFFT2D_ERR_SIZE, "eps = %g gives %d taps per side; at most %d (eps >= ~1e-14)", eps, p.msp,
                           kMaxMsp);
    const int mx = fft_len(p.Mrx), my = fft_len(p.Mry);
    if (mx > kMaxFft || my > kMaxFft)
        return nfft2d_fail(NFFT2D_ERR_SIZE,
                           "grid %d x %d (Mrx x Mry) needs FFTs of %d and %d points; at most %d (Mrx, Mry <= 4096, "
                           "or 8192 for powers of two): reduce the products of the extents X1 S and Y1 T",
                           p.Mrx, p.Mry, mx, my, kMaxFft);
    p.nbx = (p.Mrx + kTile - 1) / kTile;
    p.nby = (p.Mry + kTile - 1) / kTile;
    p.ibx = (p.Mrx + kInBin - 1) / kInBin;
    p.iby = (p.Mry + kInBin - 1) / kInBin;
    const int nb = p.nbx * p.nby;       // tiles / output bins
    const int nki = p.ibx * p.iby;      // input keys: 8 x 8-cell bins
    const int nko = nb * kTile;         // output keys: bin and the row inside it
    const size_t cells = (size_t)p.Mrx * p.Mry;

    // workspace: counts, starts, records, the grid and its transpose, FFT tables
    size_t need = 0;
    {
        const size_t sz[] = {al(sizeof(int) * (nki + 2)) * 2 + al(sizeof(int) * (nko + 2)) * 2,
                             al(sizeof(int) * ((nki + 1) / 1024 + 2)) + al(sizeof(int) * ((nko + 1) / 1024 + 2)),
                             al(sizeof(InRec) * len_in) * 2, al(sizeof(int2) * len_in) * 2,
                             al(sizeof(OutRec) * len_out) * 2, al(sizeof(int2) * len_out),
                             al(sizeof(double2) * cells) * 2,
                             al(sizeof(double2) * mx) * 2 + al(sizeof(double2) * p.Mrx),
                             al(sizeof(double2) * my) * 2 + al(sizeof(double2) * p.Mry)};
        for (size_t v : sz) need += v;
    }
    if ((rc = ensure(ws, need))) return rc;
    Carve cv{ws.ptr};
    int *cnt_in = cv.take<int>(nki + 2), *cnt_out = cv.take<int>(nko + 2);
    int *st_in = cv.take<int>(nki + 2), *st_out = cv.take<int>(nko + 2);
    int *tot_in = cv.take<int>((nki + 1) / 1024 + 2), *tot_out = cv.take<int>((nko + 1) / 1024 + 2);
    InRec *rin = cv.take<InRec>(len_in), *sin_ = cv.take<InRec>(len_in);
    int2 *kin = cv.take<int2>(len_in), *cin = cv.take<int2>(len_in);
    OutRec *rout = cv.take<OutRec>(len_out), *sout = cv.take<OutRec>(len_out);
    int2 *kout = cv.take<int2>(len_out);
    double2 *G = cv.take<double2>(cells), *GT = cv.take<double2>(cells);
    double2 *twx = cv.take<double2>(mx), *bhx = cv.take<double2>(mx), *chx = cv.take<double2>(p.Mrx);
    double2 *twy = cv.take<double2>(my), *bhy = cv.take<double2>(my), *chy = cv.take<double2>(p.Mry);

    hipStream_t st = nullptr;
    NFFT2D_HIP_TRY(hipMemsetAsync(cnt_in, 0, al(sizeof(int) * (nki + 2)) + al(sizeof(int) * (nko + 2)), st));
    const unsigned gin = (unsigned)((len_in + 255) / 256), gout = (unsigned)((len_out + 255) / 256);
    hipLaunchKernelGGL(k3_prep_in, dim3(gin), dim3(256), 0, st, x, y, reinterpret_cast<const double2 *>(f), len_in,
                       p, rin, kin, cnt_in);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k3_prep_out, dim3(gout), dim3(256), 0, st, s, t, len_out, p, rout, kout, cnt_out,
                       reinterpret_cast<double2 *>(F));
    NFFT2D_LAUNCH_CHECK();
    // keys: nki bins (+ 1 out-of-grid) for the inputs, nko bin rows (+ 1) for the outputs
    for (int side = 0; side < 2; ++side) {
        const int n = (side ? nko : nki) + 1, segs = (n + 1023) / 1024;
        int *cnt = side ? cnt_out : cnt_in, *start = side ? st_out : st_in, *tot = side ? tot_out : tot_in;
        hipLaunchKernelGGL(k3_scan_local, dim3(segs), dim3(1024), 0, st, cnt, n, start, tot);
        NFFT2D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k3_scan_totals, dim3(1), dim3(1024), 0, st, tot, segs);
        NFFT2D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k3_scan_add, dim3(n / 256 + 1), dim3(256), 0, st, start, n, tot);
        NFFT2D_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k3_scatter_in, dim3(gin), dim3(256), 0, st, rin, kin, len_in, st_in, sin_, cin);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k3_scatter<OutRec>, dim3(gout), dim3(256), 0, st, rout, kout, len_out, st_out, sout);
    NFFT2D_LAUNCH_CHECK();

    // FFT tables (twiddles; chirps and their transforms for Bluestein lines)
    auto tables = [&](int n, int m, double2 *tw, double2 *ch, double2 *bh) -> int {
        hipLaunchKernelGGL(k3_twiddles, dim3((m + 255) / 256), dim3(256), 0, st, m, tw);
        NFFT2D_LAUNCH_CHECK();
        if (is_pow2(n)) return NFFT2D_OK;
        hipLaunchKernelGGL(k3_chirp, dim3((m + 255) / 256), dim3(256), 0, st, n, m, ch, bh);
        NFFT2D_LAUNCH_CHECK();
        const int T = std::min(1024, std::max(64, m / 4));
        hipLaunchKernelGGL(k3_fft_lines, dim3(1), dim3(T), sizeof(double2) * fft_lds_entries(m), st, bh, m, m, tw,
                           (const double2 *)nullptr, (const double2 *)nullptr);
        NFFT2D_LAUNCH_CHECK();
        return NFFT2D_OK;
    };
    NFFT2D_HIP_TRY(hipFuncSetAttribute((const void *)k3_fft_lines, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sizeof(double2) * fft_lds_entries(kMaxFft)));
    if ((rc = tables(p.Mrx, mx, twx, chx, bhx))) return rc;
    if ((rc = tables(p.Mry, my, twy, chy, bhy))) return rc;

    // spread
    hipLaunchKernelGGL(k3_spread, dim3(nb), dim3(64 * kSpreadWaves), 0, st, sin_, cin, st_in, p, G);
    NFFT2D_LAUNCH_CHECK();

    // 2-D DFT: rows (length Mrx), transpose, rows of the transpose (length Mry)
    auto lines = [&](double2 *data, int n, int m, int count, const double2 *tw, const double2 *ch,
                     const double2 *bh) -> int {
        const int T = std::min(1024, std::max(64, m / 4));
        const bool blue = !is_pow2(n);
        hipLaunchKernelGGL(k3_fft_lines, dim3(count), dim3(T), sizeof(double2) * fft_lds_entries(m), st, data, n, m,
                           tw,
                           blue ? ch : (const double2 *)nullptr, blue ? bh : (const double2 *)nullptr);
        NFFT2D_LAUNCH_CHECK();
        return NFFT2D_OK;
    };
    if ((rc = lines(G, p.Mrx, mx, p.Mry, twx, chx, bhx))) return rc;
    hipLaunchKernelGGL(k3_transpose, dim3((p.Mrx + 31) / 32, (p.Mry + 31) / 32), dim3(256), 0, st, G, GT, p.Mry,
                       p.Mrx);
    NFFT2D_LAUNCH_CHECK();
    if ((rc = lines(GT, p.Mry, my, p.Mrx, twy, chy, bhy))) return rc;

    // interpolation at the outputs
    const int W = kTile + 2 * p.msp;
    const size_t lds = sizeof(double2) * (size_t)W * W;
    NFFT2D_HIP_TRY(hipFuncSetAttribute((const void *)k3_interp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k3_interp, dim3(nb), dim3(64 * kInterpWaves), lds, st, GT, sout, st_out, p,
                       reinterpret_cast<double2 *>(F));
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

}  // namespace

extern "C" int NUFFT_2D_GPU(const double *x, const double *y, const double *s, const double *t,
                            const nfft2d_c128 *f, nfft2d_c128 *F, double eps, int len_in, int len_out)
{
    return run(x, y, s, t, f, F, eps, len_in, len_out);
}

extern "C" int NUFFT_2D_GPU_params(const double *x, const double *y, const double *s, const double *t, double eps,
                                   int len_in, int len_out, double *out)
{
    int rc = check_args(x, y, s, t, (const void *)1, (const void *)1, eps, len_in, len_out);
    if (rc) return rc;
    if (!out) return nfft2d_fail(NFFT2D_ERR_ARG, "null output array");
    if (len_in == 0 || len_out == 0) return nfft2d_fail(NFFT2D_ERR_ARG, "no inputs or no outputs");
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    NFFT2D_HIP_TRY(hipGetDevice(&dev));
    return params(g_ws[dev], x, y, s, t, eps, len_in, len_out, out);
}

extern "C" int NUFFT_2D_GPU_host(const double *x, const double *y, const double *s, const double *t,
                                 const nfft2d_c128 *f, nfft2d_c128 *F, double eps, int len_in, int len_out)
{
    int rc = check_args(x, y, s, t, f, F, eps, len_in, len_out);
    if (rc) return rc;
    if (len_out == 0) return NFFT2D_OK;
    const size_t ni = (size_t)len_in, no = (size_t)len_out;
    void *buf = nullptr;
    const size_t bytes = al(16 * ni) + al(16 * ni) + al(16 * no) * 2;
    if (nfft2d_malloc(&buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "NUFFT_2D_GPU_host: %zu bytes", bytes);
    }
    unsigned char *b = static_cast<unsigned char *>(buf);
    double *dx = reinterpret_cast<double *>(b), *dy = dx + ni;
    nfft2d_c128 *df = reinterpret_cast<nfft2d_c128 *>(b + al(16 * ni));
    double *ds = reinterpret_cast<double *>(b + al(16 * ni) * 2), *dt = ds + no;
    nfft2d_c128 *dF = reinterpret_cast<nfft2d_c128 *>(b + al(16 * ni) * 2 + al(16 * no));
    hipError_t e = hipSuccess;
    if (ni) {
        if (e == hipSuccess) e = hipMemcpy(dx, x, 8 * ni, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dy, y, 8 * ni, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(df, f, 16 * ni, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(ds, s, 8 * no, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, t, 8 * no, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(buf);
        return nfft2d_fail(NFFT2D_ERR_HIP, "NUFFT_2D_GPU_host: copy in: %s", hipGetErrorString(e));
    }
    rc = run(ni ? dx : nullptr, ni ? dy : nullptr, ds, dt, ni ? df : nullptr, dF, eps, len_in, len_out);
    if (rc == NFFT2D_OK) {
        e = hipMemcpy(F, dF, 16 * no, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = nfft2d_fail(NFFT2D_ERR_HIP, "NUFFT_2D_GPU_host: copy out: %s", hipGetErrorString(e));
    }
    (void)hipFree(buf);
    return rc;
}
