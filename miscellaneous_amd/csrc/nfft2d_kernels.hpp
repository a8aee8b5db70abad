// nfft2d_kernels.hpp -- gfx950 kernels of the 2D NFFT (type-2, type-1 and the
// gradient), templated on the real type T (float / double).
//
// Grid conventions (internal; the c*N1 x c*N2 oversampled grid never leaves
// the library):
//   * the grid is kept UNSHIFTED: spatial index n in [0, L), L = 2N, and mode k
//     lives at DFT index k mod L.  That folds the reference's fftshift_2D /
//     ifftshift passes (NFFT/NFFT2_2D.cu:383,393; BP_FBP_FFBP/NFFT1_2D.m:55)
//     into index arithmetic.
//   * the grid is stored transposed, G[n2][n1] (n1 contiguous), so that the
//     pass that writes it (type-2) or reads it (type-1) streams whole rows.
//   * a 2N-point DFT whose input (type-2) or output (type-1) is zero outside
//     k in [-N/2, N/2) is computed as two N-point FFTs (even/odd outputs, or
//     even/odd inputs) with a pre/post twiddle e^{-+ pi i k / N}.
#pragma once

#include "nfft2d_device.hpp"

namespace nfft2d {

// ---------------------------------------------------------------- small helpers
template <typename C> __device__ __forceinline__ void store_pair(C* p, C a, C b);
template <> __device__ __forceinline__ void store_pair<float2>(float2* p, float2 a, float2 b)
{
    *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
}
template <> __device__ __forceinline__ void store_pair<double2>(double2* p, double2 a, double2 b)
{
    p[0] = a;
    p[1] = b;
}
template <typename C> __device__ __forceinline__ void load_pair(const C* p, C& a, C& b);
template <> __device__ __forceinline__ void load_pair<float2>(const float2* p, float2& a, float2& b)
{
    float4 q = *reinterpret_cast<const float4*>(p);
    a = make_float2(q.x, q.y);
    b = make_float2(q.z, q.w);
}
template <> __device__ __forceinline__ void load_pair<double2>(const double2* p, double2& a, double2& b)
{
    a = p[0];
    b = p[1];
}

__device__ __forceinline__ int wrap(int g, int L)
{
    g %= L;
    return g < 0 ? g + L : g;
}

// Row-group / column-block sizes of the FFT passes.
template <int N> constexpr int rows_per_block() { return N >= 1024 ? 1 : 1024 / N; }
template <int N> constexpr int cols_per_block() { return N >= 2048 ? 1 : 2048 / N; }

// ---------------------------------------------------------------- fill
template <typename T>
__global__ void k_fill_zero(T* __restrict__ p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (T)0;
}

// ================================================================= type-2 passes
// (a)+(b) row pass: T[i1][n2] = D1[i1] sum_{k2} f[i1][k2] D2[k2] e^{-2 pi i k2 n2 / L2},
// n2 in [0, L2).  The window pre-compensation and zero padding of
// BP_FBP_FFBP/NFFT1_2D.m:48-49 are applied while loading; the padding is
// never materialised.
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(rows_per_block<N>() * (N / 8)) void k_t2_rows(
    const cplx<T>* __restrict__ f, cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    using C = cplx<T>;
    constexpr int ROWS = rows_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    __shared__ C lds[ROWS * V * LL];
    const int g = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int i1 = blockIdx.x * ROWS + g;
    const int row = i1 < N1 ? i1 : N1 - 1;
    const T d1 = D1[row];
    C v[V][8];
#pragma unroll
    for (int b = 0; b < 8 / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int p = first_in_index<N>(j, b, r);
            const int k2 = p < N / 2 ? p : p - N;
            const int i2 = k2 + N / 2;
            C a = cscale(f[(size_t)row * N + i2], d1 * D2[i2]);
            v[0][b * R0 + r] = a;
            v[1][b * R0 + r] = cmul(a, twid<SIGN>(tw, k2 < 0 ? k2 + 2 * N : k2));
        }
    C* mine = lds + g * V * LL;
    fft_lines<N, V, SIGN>(v, mine, j, mine, j, tw, 2 * N);
    if (i1 < N1) {
#pragma unroll
        for (int b = 0; b < 8 / RL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int q = last_out_index<N>(j, b, r);
                store_pair(&Tm[(size_t)row * (2 * N) + 2 * q], v[0][b * RL + r], v[1][b * RL + r]);
            }
    }
}

// (b) column pass: Gt[n2][n1] = sum_{k1} T[k1][n2] e^{-2 pi i k1 n1 / L1}, n1 in [0, L1).
// Loads B columns of T per block (B consecutive elements per row), writes B
// contiguous rows of the transposed grid.
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(cols_per_block<N>() * (N / 8)) void k_t2_cols(
    const cplx<T>* __restrict__ Tm, cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int L2)
{
    using C = cplx<T>;
    constexpr int B = cols_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    __shared__ C lds[B * V * LL];
    const int t = threadIdx.x;
    const int bF = t % B, jF = t / B;      // first stage: B columns side by side
    const int bL = t / TPL, jL = t % TPL;  // last stage: whole lines
    const int n2F = min(blockIdx.x * B + bF, L2 - 1);
    C v[V][8];
#pragma unroll
    for (int b = 0; b < 8 / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int p = first_in_index<N>(jF, b, r);
            const int k1 = p < N / 2 ? p : p - N;
            C a = Tm[(size_t)(k1 + N / 2) * L2 + n2F];
            v[0][b * R0 + r] = a;
            v[1][b * R0 + r] = cmul(a, twid<SIGN>(tw, k1 < 0 ? k1 + 2 * N : k1));
        }
    fft_lines<N, V, SIGN>(v, lds + bF * V * LL, jF, lds + bL * V * LL, jL, tw, 2 * N);
    const int n2L = blockIdx.x * B + bL;
    if (n2L < L2) {
        C* out = Gt + (size_t)n2L * (2 * N);
#pragma unroll
        for (int b = 0; b < 8 / RL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int q = last_out_index<N>(jL, b, r);
                store_pair(&out[2 * q], v[0][b * RL + r], v[1][b * RL + r]);
            }
    }
}

// ================================================================= type-1 passes
// Column pass: Tp[k1][n2] = sum_{n1} Gt[n2][n1] e^{SIGN 2 pi i n1 k1 / L1} for
// k1 in [-N1/2, N1/2) only (the reference crops the rest, NFFT/NFFT2_2D.cu:74-86).
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(cols_per_block<N>() * (N / 8)) void k_t1_cols(
    const cplx<T>* __restrict__ Gt, cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int L2)
{
    using C = cplx<T>;
    constexpr int B = cols_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    __shared__ C lds[B * V * LL];
    const int t = threadIdx.x;
    const int bF = t / TPL, jF = t % TPL;  // first stage: whole grid rows (contiguous)
    const int bL = t % B, jL = t / B;      // last stage: B outputs side by side
    const int n2F = min(blockIdx.x * B + bF, L2 - 1);
    const C* in = Gt + (size_t)n2F * (2 * N);
    C v[V][8];
#pragma unroll
    for (int b = 0; b < 8 / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int p = first_in_index<N>(jF, b, r);
            load_pair(&in[2 * p], v[0][b * R0 + r], v[1][b * R0 + r]);
        }
    fft_lines<N, V, SIGN>(v, lds + bF * V * LL, jF, lds + bL * V * LL, jL, tw, 2 * N);
    const int n2L = blockIdx.x * B + bL;
    if (n2L < L2) {
#pragma unroll
        for (int b = 0; b < 8 / RL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int q = last_out_index<N>(jL, b, r);
                const int k1 = q < N / 2 ? q : q - N;
                C y = cadd(v[0][b * RL + r], cmul(v[1][b * RL + r], twid<SIGN>(tw, k1 < 0 ? k1 + 2 * N : k1)));
                Tp[(size_t)(k1 + N / 2) * L2 + n2L] = y;
            }
    }
}

// Row pass + crop + deconvolution: F[k1][k2] = D1[k1] D2[k2] sum_{n2} Tp[k1][n2] e^{SIGN 2 pi i n2 k2 / L2}
// (NFFT/NFFT2_2D.cu:74-100 Decimation_and_Scaling fused into the last FFT stage).
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(rows_per_block<N>() * (N / 8)) void k_t1_rows(
    const cplx<T>* __restrict__ Tp, cplx<T>* __restrict__ F, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    using C = cplx<T>;
    constexpr int ROWS = rows_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    __shared__ C lds[ROWS * V * LL];
    const int g = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int i1 = blockIdx.x * ROWS + g;
    const int row = i1 < N1 ? i1 : N1 - 1;
    const C* in = Tp + (size_t)row * (2 * N);
    C v[V][8];
#pragma unroll
    for (int b = 0; b < 8 / R0; ++b)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int p = first_in_index<N>(j, b, r);
            load_pair(&in[2 * p], v[0][b * R0 + r], v[1][b * R0 + r]);
        }
    C* mine = lds + g * V * LL;
    fft_lines<N, V, SIGN>(v, mine, j, mine, j, tw, 2 * N);
    if (i1 < N1) {
        const T d1 = D1[row];
#pragma unroll
        for (int b = 0; b < 8 / RL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int q = last_out_index<N>(j, b, r);
                const int k2 = q < N / 2 ? q : q - N;
                C y = cadd(v[0][b * RL + r], cmul(v[1][b * RL + r], twid<SIGN>(tw, k2 < 0 ? k2 + 2 * N : k2)));
                F[(size_t)row * N + (k2 + N / 2)] = cscale(y, d1 * D2[k2 + N / 2]);
            }
    }
}

// ================================================================= direct-DFT passes
// Same four passes for any even N (not a power of two, or N < 8): each output
// is a direct sum over the line, staged in LDS, with twiddle index
// (k n) mod L into the 2N-entry table.  O(N^2) per line; used for the
// reference's non-power-of-two sizes (e.g. 2Nu = 88 in Synthesis/PSO).
template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t2_rows_direct(
    const cplx<T>* __restrict__ f, cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1, int N2)
{
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* line = reinterpret_cast<C*>(smem_raw);
    const int i1 = blockIdx.x, L2 = 2 * N2;
    for (int i2 = threadIdx.x; i2 < N2; i2 += blockDim.x)
        line[i2] = cscale(f[(size_t)i1 * N2 + i2], D1[i1] * D2[i2]);
    __syncthreads();
    for (int n2 = threadIdx.x; n2 < L2; n2 += blockDim.x) {
        C acc = mk((T)0, (T)0);
        for (int i2 = 0; i2 < N2; ++i2) {
            const int k2 = i2 - N2 / 2;
            const long long e = ((long long)(k2 < 0 ? k2 + L2 : k2) * n2) % L2;
            acc = cadd(acc, cmul(line[i2], twid<SIGN>(tw, (int)e)));
        }
        Tm[(size_t)i1 * L2 + n2] = acc;
    }
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t2_cols_direct(
    const cplx<T>* __restrict__ Tm, cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int N1, int L2)
{
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* line = reinterpret_cast<C*>(smem_raw);
    const int n2 = blockIdx.x, L1 = 2 * N1;
    for (int i1 = threadIdx.x; i1 < N1; i1 += blockDim.x) line[i1] = Tm[(size_t)i1 * L2 + n2];
    __syncthreads();
    for (int n1 = threadIdx.x; n1 < L1; n1 += blockDim.x) {
        C acc = mk((T)0, (T)0);
        for (int i1 = 0; i1 < N1; ++i1) {
            const int k1 = i1 - N1 / 2;
            const long long e = ((long long)(k1 < 0 ? k1 + L1 : k1) * n1) % L1;
            acc = cadd(acc, cmul(line[i1], twid<SIGN>(tw, (int)e)));
        }
        Gt[(size_t)n2 * L1 + n1] = acc;
    }
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t1_cols_direct(
    const cplx<T>* __restrict__ Gt, cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int N1, int L2)
{
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* line = reinterpret_cast<C*>(smem_raw);
    const int n2 = blockIdx.x, L1 = 2 * N1;
    for (int n1 = threadIdx.x; n1 < L1; n1 += blockDim.x) line[n1] = Gt[(size_t)n2 * L1 + n1];
    __syncthreads();
    for (int i1 = threadIdx.x; i1 < N1; i1 += blockDim.x) {
        const int k1 = i1 - N1 / 2;
        const long long kk = k1 < 0 ? k1 + L1 : k1;
        C acc = mk((T)0, (T)0);
        for (int n1 = 0; n1 < L1; ++n1) acc = cadd(acc, cmul(line[n1], twid<SIGN>(tw, (int)((kk * n1) % L1))));
        Tp[(size_t)i1 * L2 + n2] = acc;
    }
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t1_rows_direct(
    const cplx<T>* __restrict__ Tp, cplx<T>* __restrict__ F, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1, int N2)
{
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* line = reinterpret_cast<C*>(smem_raw);
    const int i1 = blockIdx.x, L2 = 2 * N2;
    for (int n2 = threadIdx.x; n2 < L2; n2 += blockDim.x) line[n2] = Tp[(size_t)i1 * L2 + n2];
    __syncthreads();
    for (int i2 = threadIdx.x; i2 < N2; i2 += blockDim.x) {
        const int k2 = i2 - N2 / 2;
        const long long kk = k2 < 0 ? k2 + L2 : k2;
        C acc = mk((T)0, (T)0);
        for (int n2 = 0; n2 < L2; ++n2) acc = cadd(acc, cmul(line[n2], twid<SIGN>(tw, (int)((kk * n2) % L2))));
        F[(size_t)i1 * N2 + i2] = cscale(acc, D1[i1] * D2[i2]);
    }
}

// ================================================================= setpts (bin sort)
// Points are counting-sorted by bin: a bin is a kBin x kBin block of cells of
// the oversampled grid, holding mu = rint(c x) mod L.  Bins are the tiles of the
// gather and the output tiles of the spread.
constexpr int kBin = 16;
constexpr int kMaxLdsBins = 16384;      // per-block LDS histogram limit (64 KiB)

template <typename T>
__device__ __forceinline__ int bin_of(T x, T y, int L1, int L2, T inv1, T inv2, int nb1, int& m1, int& m2, T& t1,
                                      T& t2)
{
    locate(x, L1, inv1, m1, t1);
    locate(y, L2, inv2, m2, t2);
    return (m2 / kBin) * nb1 + m1 / kBin;
}

// Pass 1: block b counts the bins of points [b*chunk, (b+1)*chunk) in an LDS
// histogram (integer LDS atomics); rank[i] is the point's rank inside its
// (block, bin).  hist is [nblk][nbins].
template <typename T>
__global__ __launch_bounds__(256) void k_hist(const T* __restrict__ x, const T* __restrict__ y, int M, int chunk,
                                              int L1, int L2, T inv1, T inv2, int nb1, int nbins,
                                              int* __restrict__ hist, int* __restrict__ rank)
{
    extern __shared__ int h[];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int lo = blockIdx.x * chunk, hi = min(lo + chunk, M);
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        int m1, m2;
        T t1, t2;
        const int bin = bin_of(x[i], y[i], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
        rank[i] = atomicAdd(&h[bin], 1);
    }
    __syncthreads();
    int* out = hist + (size_t)blockIdx.x * nbins;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) out[i] = h[i];
}

// Pass 1 for grids with more than kMaxLdsBins bins: one logical block, global
// integer atomics.
template <typename T>
__global__ __launch_bounds__(256) void k_hist_global(const T* __restrict__ x, const T* __restrict__ y, int M, int L1,
                                                     int L2, T inv1, T inv2, int nb1, int* __restrict__ hist,
                                                     int* __restrict__ rank)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int m1, m2;
    T t1, t2;
    const int bin = bin_of(x[i], y[i], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
    rank[i] = atomicAdd(&hist[bin], 1);
}

// Pass 2: per bin, turn hist[blk][bin] into the offset of (blk, bin) inside
// the bin and write the bin total.  A block covers 32 bins x 8 slices of the
// block range, so the loads of a slice are independent and coalesced.
__global__ __launch_bounds__(256) void k_colscan(int* __restrict__ hist, int nblk, int nbins, int* __restrict__ tot)
{
    __shared__ int part[8][33];
    const int bl = threadIdx.x % 32, q = threadIdx.x / 32;
    const int bin = blockIdx.x * 32 + bl;
    const int per = (nblk + 7) / 8;
    const int lo = min(q * per, nblk), hi = min(lo + per, nblk);
    int sum = 0;
    if (bin < nbins)
        for (int k = lo; k < hi; ++k) sum += hist[(size_t)k * nbins + bin];
    part[q][bl] = sum;
    __syncthreads();
    if (q == 0) {
        int run = 0;
        for (int j = 0; j < 8; ++j) {
            const int v = part[j][bl];
            part[j][bl] = run;
            run += v;
        }
        if (bin < nbins) tot[bin] = run;
    }
    __syncthreads();
    int run = part[q][bl];
    if (bin < nbins)
        for (int k = lo; k < hi; ++k) {
            const int v = hist[(size_t)k * nbins + bin];
            hist[(size_t)k * nbins + bin] = run;
            run += v;
        }
}

// Pass 3 (one block): exclusive scan of the bin totals -> binstart, and the
// subproblem list (bin, start, count <= maxpts) in bin order.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_scan_bins(const int* __restrict__ tot, int nbins, int maxpts,
                                                            int* __restrict__ binstart, int4* __restrict__ sub,
                                                            int* __restrict__ nsub)
{
    __shared__ int w_pts[kScanThreads / 64], w_sub[kScanThreads / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (nbins + kScanThreads - 1) / kScanThreads;
    const int lo = min(t * per, nbins), hi = min(lo + per, nbins);
    int a = 0, b = 0;
    for (int i = lo; i < hi; ++i) {
        const int c = tot[i];
        a += c;
        b += (c + maxpts - 1) / maxpts;
    }
    int ia = a, ib = b;   // inclusive wave scans
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int na = __shfl_up(ia, o, 64), nb = __shfl_up(ib, o, 64);
        if (lane >= o) { ia += na; ib += nb; }
    }
    if (lane == 63) { w_pts[w] = ia; w_sub[w] = ib; }
    __syncthreads();
    int ba = 0, bb = 0;
    for (int k = 0; k < w; ++k) { ba += w_pts[k]; bb += w_sub[k]; }
    int pos = ba + ia - a, sp = bb + ib - b;
    for (int i = lo; i < hi; ++i) {
        const int c = tot[i];
        binstart[i] = pos;
        for (int k = 0; k * maxpts < c; ++k) sub[sp++] = make_int4(i, pos + k * maxpts, min(maxpts, c - k * maxpts), 0);
        pos += c;
    }
    if (t == kScanThreads - 1) *nsub = sp;
}

// Pass 5: scatter the points to their sorted positions.
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const T* __restrict__ x, const T* __restrict__ y, int M, int chunk,
                                                 int L1, int L2, T inv1, T inv2, int nb1, int nbins,
                                                 const int* __restrict__ off, const int* __restrict__ binstart,
                                                 const int* __restrict__ rank, cplx<T>* __restrict__ xy,
                                                 int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const T xi = x[i], yi = y[i];
    int m1, m2;
    T t1, t2;
    const int bin = bin_of(xi, yi, L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
    const int p = binstart[bin] + off[(size_t)(i / chunk) * nbins + bin] + rank[i];
    xy[p] = mk(xi, yi);
    idx[p] = i;
}

// ================================================================= window taps
struct WinArgs {
    WinPoly val;    // phi_hat / S
    WinPoly der;    // phi_hat' / S
    WinExact ex;    // FP64 closed form
};

__device__ __forceinline__ void taps(float t, const WinArgs& wa, float (&w)[kW]) { window_taps(t, wa.val, w); }
__device__ __forceinline__ void taps(double t, const WinArgs& wa, double (&w)[kW])
{
#pragma unroll
    for (int m = 0; m < kW; ++m) w[m] = window_exact(t - (double)(m - kK), wa.ex);
}

// Derivative taps d/dt phi_hat(t - (m - K)) / S.  phi_hat' is odd, so tap -j
// is -E'_j(t^2) + t O'_j(t^2); for j = 0 only the odd part exists.
__device__ __forceinline__ void dtaps(float t, const WinArgs& wa, float (&w)[kW])
{
    const float u = t * t;
#pragma unroll
    for (int j = 0; j <= kK; ++j) {
        float o = wa.der.o[j][kNO - 1];
#pragma unroll
        for (int q = kNO - 2; q >= 0; --q) o = fmaf(o, u, wa.der.o[j][q]);
        if (j == 0) {
            w[kK] = t * o;
        } else {
            float e = wa.der.e[j][kNE - 1];
#pragma unroll
            for (int q = kNE - 2; q >= 0; --q) e = fmaf(e, u, wa.der.e[j][q]);
            w[kK + j] = fmaf(t, o, e);
            w[kK - j] = fmaf(t, o, -e);
        }
    }
}
__device__ __forceinline__ void dtaps(double t, const WinArgs& wa, double (&w)[kW])
{
#pragma unroll
    for (int m = 0; m < kW; ++m) w[m] = window_exact_deriv(t - (double)(m - kK), wa.ex);
}

// Geometry shared by the gather and spread kernels.
struct BinGeom {
    int L1, L2, nb1;
    float inv1f, inv2f;
    double inv1d, inv2d;
};
__device__ __forceinline__ float inv1(const BinGeom& g, float) { return g.inv1f; }
__device__ __forceinline__ float inv2(const BinGeom& g, float) { return g.inv2f; }
__device__ __forceinline__ double inv1(const BinGeom& g, double) { return g.inv1d; }
__device__ __forceinline__ double inv2(const BinGeom& g, double) { return g.inv2d; }

// Load the (kBin+2K)^2 region of the transposed grid around bin (b1, b2) into LDS.
// o1, o2 are in [-K, L); when L >= W one adjustment by L wraps every index.
__device__ __forceinline__ int wrap1(int g, int L) { return g < 0 ? g + L : (g >= L ? g - L : g); }

template <typename C>
__device__ __forceinline__ void load_tile(C* tile, const C* __restrict__ Gt, int L1, int L2, int o1, int o2)
{
    constexpr int W = kBin + 2 * kK;
    if (L1 >= W && L2 >= W) {
        for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
            const int r = e / W, c = e - r * W;
            tile[e] = Gt[(size_t)wrap1(o2 + r, L2) * L1 + wrap1(o1 + c, L1)];
        }
    } else {
        for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
            const int r = e / W, c = e - r * W;
            tile[e] = Gt[(size_t)wrap(o2 + r, L2) * L1 + wrap(o1 + c, L1)];
        }
    }
}

constexpr int kGatherThreads = 256;     // = points per gather subproblem

// Exclusive scan of kGatherThreads ints held one per thread (block-wide).
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_warp)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int n = __shfl_up(inc, o, 64);
        if (lane >= o) inc += n;
    }
    if (lane == 63) s_warp[w] = inc;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += s_warp[k];
    __syncthreads();
    return base + inc - v;
}

// (c) type-2 gather over one subproblem (<= 256 points of one bin), MODE 0:
// out[l] = y_l;  MODE 1: out = y, out_d1 = dy/dx1, out_d2 = dy/dx2;  MODE 2:
// gradient Re(conj(g_l) dy_l/dx) into gx1/gx2.
// y_l = sum_{m2} w2[m2] sum_{m1} w1[m1] G[(mu2+m2-K) mod L2][(mu1+m1-K) mod L1]
// (BP_FBP_FFBP/NFFT1_2D.m:61-67 on the unshifted grid).  The block first
// counting-sorts its points by cell in LDS so that neighbouring lanes read
// neighbouring tile addresses.
template <typename T, int MODE>
__global__ __launch_bounds__(kGatherThreads) void k_gather(
    const cplx<T>* __restrict__ Gt, BinGeom geo, const cplx<T>* __restrict__ xy, const int* __restrict__ idx,
    const int4* __restrict__ sub, const int* __restrict__ nsub, WinArgs wa, cplx<T>* __restrict__ out,
    cplx<T>* __restrict__ out_d1, cplx<T>* __restrict__ out_d2, const cplx<T>* __restrict__ g, T* __restrict__ gx1,
    T* __restrict__ gx2)
{
    using C = cplx<T>;
    constexpr int W = kBin + 2 * kK;
    __shared__ C tile[W * W];
    __shared__ int cnt[kBin * kBin];
    __shared__ int s_warp[kGatherThreads / 64];
    __shared__ T s_f1[kGatherThreads], s_f2[kGatherThreads];
    __shared__ int s_cell[kGatherThreads], s_idx[kGatherThreads];
    const int s = blockIdx.x;
    if (s >= *nsub) return;
    const int4 d = sub[s];
    const int b1 = d.x % geo.nb1, b2 = d.x / geo.nb1;
    const int O1 = b1 * kBin, O2 = b2 * kBin;
    const int tid = threadIdx.x;
    const bool have = tid < d.z;
    int cell = 0, o = 0;
    T f1 = 0, f2 = 0;
    if (have) {
        const C q = xy[d.y + tid];
        o = idx[d.y + tid];
        int m1, m2;
        locate(q.x, geo.L1, inv1(geo, T()), m1, f1);
        locate(q.y, geo.L2, inv2(geo, T()), m2, f2);
        cell = min(max(m2 - O2, 0), kBin - 1) * kBin + min(max(m1 - O1, 0), kBin - 1);
    }
    cnt[tid] = 0;
    load_tile(tile, Gt, geo.L1, geo.L2, O1 - kK, O2 - kK);
    __syncthreads();
    const int rk = have ? atomicAdd(&cnt[cell], 1) : 0;
    __syncthreads();
    const int start = block_exclusive_scan(cnt[tid], s_warp);
    __syncthreads();
    cnt[tid] = start;
    __syncthreads();
    if (have) {
        const int dst = cnt[cell] + rk;
        s_f1[dst] = f1;
        s_f2[dst] = f2;
        s_cell[dst] = cell;
        s_idx[dst] = o;
    }
    __syncthreads();
    if (!have) return;
    f1 = s_f1[tid];
    f2 = s_f2[tid];
    cell = s_cell[tid];
    o = s_idx[tid];
    const int r0 = cell / kBin, c0 = cell % kBin;
    T w1[kW], w2[kW];
    taps(f1, wa, w1);
    taps(f2, wa, w2);
    const C* base = tile + r0 * W + c0;
    if constexpr (MODE == 0) {
        T ax = 0, ay = 0;
#pragma unroll
        for (int b = 0; b < kW; ++b) {
            T sx = 0, sy = 0;
#pragma unroll
            for (int a = 0; a < kW; ++a) {
                const C u = base[b * W + a];
                sx = fma(w1[a], u.x, sx);
                sy = fma(w1[a], u.y, sy);
            }
            ax = fma(w2[b], sx, ax);
            ay = fma(w2[b], sy, ay);
        }
        out[o] = mk(ax, ay);
    } else {
        T v1[kW], v2[kW];
        dtaps(f1, wa, v1);
        dtaps(f2, wa, v2);
        T ax = 0, ay = 0, bx = 0, by = 0, cx = 0, cy = 0;
#pragma unroll
        for (int b = 0; b < kW; ++b) {
            T sx = 0, sy = 0, dx = 0, dy = 0;
#pragma unroll
            for (int a = 0; a < kW; ++a) {
                const C u = base[b * W + a];
                sx = fma(w1[a], u.x, sx);
                sy = fma(w1[a], u.y, sy);
                dx = fma(v1[a], u.x, dx);
                dy = fma(v1[a], u.y, dy);
            }
            ax = fma(w2[b], sx, ax);
            ay = fma(w2[b], sy, ay);
            bx = fma(w2[b], dx, bx);   // d/dx1
            by = fma(w2[b], dy, by);
            cx = fma(v2[b], sx, cx);   // d/dx2
            cy = fma(v2[b], sy, cy);
        }
        const T cc = (T)kC;            // d(c x - mu)/dx = c
        if constexpr (MODE == 1) {
            out[o] = mk(ax, ay);
            out_d1[o] = mk(cc * bx, cc * by);
            out_d2[o] = mk(cc * cx, cc * cy);
        } else {
            const C gg = g[o];
            gx1[o] = cc * (gg.x * bx + gg.y * by);
            gx2[o] = cc * (gg.x * cx + gg.y * cy);
        }
    }
}

// (d) type-1 spread (interim form): each block owns one subproblem, accumulates
// the bin + K-halo in LDS, then adds it into the grid with global atomics.
template <typename T>
__global__ __launch_bounds__(256) void k_spread(cplx<T>* __restrict__ Gt, BinGeom geo, const cplx<T>* __restrict__ xy,
                                                const int* __restrict__ idx, const cplx<T>* __restrict__ data,
                                                const int4* __restrict__ sub, const int* __restrict__ nsub, WinArgs wa)
{
    using C = cplx<T>;
    constexpr int W = kBin + 2 * kK;
    __shared__ T tre[W * W];
    __shared__ T tim[W * W];
    const int s = blockIdx.x;
    if (s >= *nsub) return;
    const int4 d = sub[s];
    const int b1 = d.x % geo.nb1, b2 = d.x / geo.nb1;
    const int O1 = b1 * kBin, O2 = b2 * kBin;
    for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
        tre[e] = 0;
        tim[e] = 0;
    }
    __syncthreads();
    for (int p = d.y + threadIdx.x; p < d.y + d.z; p += blockDim.x) {
        const C q = xy[p];
        const C v = data[idx[p]];
        int m1, m2;
        T f1, f2;
        locate(q.x, geo.L1, inv1(geo, T()), m1, f1);
        locate(q.y, geo.L2, inv2(geo, T()), m2, f2);
        const int c0 = min(max(m1 - O1, 0), kBin - 1);
        const int r0 = min(max(m2 - O2, 0), kBin - 1);
        T w1[kW], w2[kW];
        taps(f1, wa, w1);
        taps(f2, wa, w2);
        T* bre = tre + r0 * W + c0;
        T* bim = tim + r0 * W + c0;
#pragma unroll
        for (int b = 0; b < kW; ++b) {
            const T vx = v.x * w2[b], vy = v.y * w2[b];
#pragma unroll
            for (int a = 0; a < kW; ++a) {
                atomicAdd(&bre[b * W + a], vx * w1[a]);
                atomicAdd(&bim[b * W + a], vy * w1[a]);
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
        const T a = tre[e], b = tim[e];
        if (a != (T)0 || b != (T)0) {
            const int r = e / W, c = e - r * W;
            C* cellp = &Gt[(size_t)wrap(O2 - kK + r, geo.L2) * geo.L1 + wrap(O1 - kK + c, geo.L1)];
            unsafeAtomicAdd(&cellp->x, a);
            unsafeAtomicAdd(&cellp->y, b);
        }
    }
}

// ================================================================= MFMA spread
// (d) type-1 spread, output-stationary: one unit per 16x16 output bin (more
// units for dense bins) computes
//     G[O2+i][O1+j] = sum_p (d_p w2_p[i + O2 - mu2_p + K]) * w1_p[j + O1 - mu1_p + K]
// over the points p of the 3x3 neighbouring bins whose footprint meets the
// bin.  Each point is an outer product of two 16-vectors; v_mfma_f32_16x16x4f32
// adds four of them per instruction into a 16x16 accumulator held in
// registers.  No atomics: every cell is written once, by its bin's unit, or by
// the merge kernel for bins split into several units.  Needs L >= 44 and at
// least 3 bins per dimension (nfft2d_plan chooses the LDS-atomic k_spread
// otherwise).
constexpr int kSpreadWaves = 4;
constexpr int kSpreadCand = 4096;      // candidates (3x3-bin points) per unit
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrap_signed(int v, int L)
{
    // v in (-L, L) -> representative in [-L/2, L/2)
    if (v >= L / 2) v -= L;
    if (v < -(L / 2)) v += L;
    return v;
}

// One block: per output bin, candidates = sum of the 3x3 neighbour bin counts;
// units = max(1, ceil(cand / kSpreadCand)); bins with several units get
// partial slots and a merge entry.
__global__ __launch_bounds__(kScanThreads) void k_spread_units(const int* __restrict__ tot, int nb1, int nb2,
                                                               int4* __restrict__ units, int* __restrict__ nunits,
                                                               int4* __restrict__ merges, int* __restrict__ nmerge)
{
    __shared__ int w_u[kScanThreads / 64], w_s[kScanThreads / 64], w_m[kScanThreads / 64];
    const int nbins = nb1 * nb2;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (nbins + kScanThreads - 1) / kScanThreads;
    const int lo = min(t * per, nbins), hi = min(lo + per, nbins);
    auto cand_of = [&](int b) {
        const int b1 = b % nb1, b2 = b / nb1;
        int c = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int y2 = (b2 + dy + nb2) % nb2, x1 = (b1 + dx + nb1) % nb1;
                c += tot[y2 * nb1 + x1];
            }
        return c;
    };
    int cu = 0, cs = 0, cm = 0;
    for (int b = lo; b < hi; ++b) {
        const int c = cand_of(b);
        const int nu = max(1, (c + kSpreadCand - 1) / kSpreadCand);
        cu += nu;
        if (nu > 1) { cs += nu; cm += 1; }
    }
    int iu = cu, is = cs, im = cm;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int a = __shfl_up(iu, o, 64), b = __shfl_up(is, o, 64), c = __shfl_up(im, o, 64);
        if (lane >= o) { iu += a; is += b; im += c; }
    }
    if (lane == 63) { w_u[w] = iu; w_s[w] = is; w_m[w] = im; }
    __syncthreads();
    int bu = 0, bs = 0, bm = 0;
    for (int k = 0; k < w; ++k) { bu += w_u[k]; bs += w_s[k]; bm += w_m[k]; }
    int uo = bu + iu - cu, so = bs + is - cs, mo = bm + im - cm;
    for (int b = lo; b < hi; ++b) {
        const int c = cand_of(b);
        const int nu = max(1, (c + kSpreadCand - 1) / kSpreadCand);
        for (int k = 0; k < nu; ++k)
            units[uo + k] = make_int4(b, k * kSpreadCand, min(c, (k + 1) * kSpreadCand), nu > 1 ? so + k : -1);
        if (nu > 1) {
            merges[mo++] = make_int4(b, so, nu, 0);
            so += nu;
        }
        uo += nu;
    }
    if (t == kScanThreads - 1) { *nunits = uo; *nmerge = mo; }
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kSpreadWaves * 64) void k_spread_mfma(
    const float2* __restrict__ xy, const int* __restrict__ idx, const float2* __restrict__ data,
    const int* __restrict__ binstart, const int* __restrict__ tot, const int4* __restrict__ units,
    const int* __restrict__ nunits, BinGeom geo, int nb2, WinArgs wa, float2* __restrict__ Gt,
    float2* __restrict__ scratch)
{
    __shared__ __attribute__((aligned(16))) float stage[kSpreadWaves][64][48];
    __shared__ int list[kSpreadWaves][128];
    const int u = blockIdx.x;
    if (u >= *nunits) return;
    const int4 U = units[u];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nb1 = geo.nb1;
    const int b1 = U.x % nb1, b2 = U.x / nb1;
    const int O1 = b1 * kBin, O2 = b2 * kBin;
    int nstart[9], npref[10];
    npref[0] = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int y2 = (b2 + j / 3 - 1 + nb2) % nb2, x1 = (b1 + j % 3 - 1 + nb1) % nb1;
        const int nb = y2 * nb1 + x1;
        nstart[j] = binstart[nb];
        npref[j + 1] = npref[j] + tot[nb];
    }
    float* st = &stage[w][0][0];
    int* lst = list[w];
    f4v acc_re = {0.f, 0.f, 0.f, 0.f}, acc_im = {0.f, 0.f, 0.f, 0.f};

    // stage k (<= 64) listed points: rows 0-15 d.x*w2, 16-31 d.y*w2, 32-47 w1 at the bin's offsets
    auto process = [&](int k) {
        float* row = st + lane * 48;
#pragma unroll
        for (int q = 0; q < 12; ++q) *reinterpret_cast<float4*>(row + 4 * q) = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < k) {
            const int p = lst[lane];
            const float2 q = xy[p];
            const float2 d = data[idx[p]];
            int m1, m2;
            float f1, f2;
            locate(q.x, geo.L1, geo.inv1f, m1, f1);
            locate(q.y, geo.L2, geo.inv2f, m2, f2);
            const int dc = wrap_signed(m1 - O1, geo.L1), dr = wrap_signed(m2 - O2, geo.L2);
            float w1[kW], w2[kW];
            taps(f1, wa, w1);
            taps(f2, wa, w2);
#pragma unroll
            for (int t = 0; t < kW; ++t) {
                const int r = dr - kK + t;
                if (r >= 0 && r < 16) {
                    row[r] = d.x * w2[t];
                    row[16 + r] = d.y * w2[t];
                }
                const int c = dc - kK + t;
                if (c >= 0 && c < 16) row[32 + c] = w1[t];
            }
        }
        wave_sync_lds();
        const int i = lane & 15, kk = lane >> 4;
        for (int s = 0; s < k; s += 4) {
            const float* pr = st + (s + kk) * 48;
            const float are = pr[i], aim = pr[16 + i], bb = pr[32 + i];
            acc_re = __builtin_amdgcn_mfma_f32_16x16x4f32(are, bb, acc_re, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, bb, acc_im, 0, 0, 0);
        }
        wave_sync_lds();
    };

    int n = 0;
    for (int c = U.y + w * 64; c < U.z; c += kSpreadWaves * 64) {
        const int cc = c + lane;
        bool rel = false;
        int p = 0;
        if (cc < U.z) {
            int j = 0;
#pragma unroll
            for (int q = 1; q < 9; ++q) j += (npref[q] <= cc) ? 1 : 0;
            p = nstart[j] + (cc - npref[j]);
            const float2 q = xy[p];
            int m1, m2;
            float f1, f2;
            locate(q.x, geo.L1, geo.inv1f, m1, f1);
            locate(q.y, geo.L2, geo.inv2f, m2, f2);
            const int dc = wrap_signed(m1 - O1, geo.L1), dr = wrap_signed(m2 - O2, geo.L2);
            rel = dr >= -kK && dr < 16 + kK && dc >= -kK && dc < 16 + kK;
        }
        const unsigned long long mask = __ballot(rel);
        if (rel) lst[n + __popcll(mask & ((1ull << lane) - 1ull))] = p;
        n += __popcll(mask);
        if (n >= 64) {
            wave_sync_lds();
            process(64);
            n -= 64;
            const int carry = lane < n ? lst[64 + lane] : 0;
            wave_sync_lds();
            if (lane < n) lst[lane] = carry;
            wave_sync_lds();
        }
    }
    if (n > 0) {
        wave_sync_lds();
        process(n);
    }
    // reduce the waves' accumulators in a fixed order
    __syncthreads();
    float* red = &stage[0][0][0];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        red[(w * 64 + lane) * 8 + r] = acc_re[r];
        red[(w * 64 + lane) * 8 + 4 + r] = acc_im[r];
    }
    __syncthreads();
    if (w == 0) {
        float vr[4], vi[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            vr[r] = red[lane * 8 + r];
            vi[r] = red[lane * 8 + 4 + r];
        }
        for (int ww = 1; ww < kSpreadWaves; ++ww)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                vr[r] += red[(ww * 64 + lane) * 8 + r];
                vi[r] += red[(ww * 64 + lane) * 8 + 4 + r];
            }
        const int j = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 4 * (lane >> 4) + r;
            if (U.w >= 0) {
                scratch[(size_t)U.w * 256 + i * 16 + j] = make_float2(vr[r], vi[r]);
            } else if (O2 + i < geo.L2 && O1 + j < geo.L1) {
                Gt[(size_t)(O2 + i) * geo.L1 + (O1 + j)] = make_float2(vr[r], vi[r]);
            }
        }
    }
}

// Sum the partial tiles of the bins split into several units, in unit order.
__global__ __launch_bounds__(256) void k_spread_merge(const int4* __restrict__ merges, const int* __restrict__ nmerge,
                                                      const float2* __restrict__ scratch, BinGeom geo,
                                                      float2* __restrict__ Gt)
{
    const int m = blockIdx.x;
    if (m >= *nmerge) return;
    const int4 M4 = merges[m];
    const int i = threadIdx.x / 16, j = threadIdx.x % 16;
    const int O1 = (M4.x % geo.nb1) * kBin, O2 = (M4.x / geo.nb1) * kBin;
    float sr = 0.f, si = 0.f;
    for (int k = 0; k < M4.z; ++k) {
        const float2 v = scratch[(size_t)(M4.y + k) * 256 + threadIdx.x];
        sr += v.x;
        si += v.y;
    }
    if (O2 + i < geo.L2 && O1 + j < geo.L1) Gt[(size_t)(O2 + i) * geo.L1 + (O1 + j)] = make_float2(sr, si);
}

}  // namespace nfft2d
