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

// ================================================================= setpts (tile sort)
// Bin = TSxTS tile of the oversampled grid that holds mu = rint(c x) mod L.
template <typename T>
__global__ __launch_bounds__(256) void k_bin_count(const T* __restrict__ x, const T* __restrict__ y, int M,
                                                   int L1, int L2, int TS, int tiles1, int* __restrict__ counts,
                                                   int* __restrict__ rank)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int m1, m2;
    T t1, t2;
    locate(x[i], L1, m1, t1);
    locate(y[i], L2, m2, t2);
    const int bin = (m2 / TS) * tiles1 + m1 / TS;
    rank[i] = atomicAdd(&counts[bin], 1);
}

// One block: exclusive scan of the tile counts and of the per-tile subproblem
// counts ceil(count / maxpts); writes the subproblem list (tile, start, count).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_scan_subprob(const int* __restrict__ counts, int ntiles,
                                                               int maxpts, int* __restrict__ offsets,
                                                               int4* __restrict__ sub, int* __restrict__ nsub)
{
    __shared__ int s_pts[kScanThreads];
    __shared__ int s_sub[kScanThreads];
    const int t = threadIdx.x;
    const int per = (ntiles + kScanThreads - 1) / kScanThreads;
    const int lo = min(t * per, ntiles), hi = min(lo + per, ntiles);
    int a = 0, b = 0;
    for (int i = lo; i < hi; ++i) {
        a += counts[i];
        b += (counts[i] + maxpts - 1) / maxpts;
    }
    s_pts[t] = a;
    s_sub[t] = b;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {   // Hillis-Steele inclusive scan
        int va = t >= off ? s_pts[t - off] : 0;
        int vb = t >= off ? s_sub[t - off] : 0;
        __syncthreads();
        s_pts[t] += va;
        s_sub[t] += vb;
        __syncthreads();
    }
    int pos = s_pts[t] - a, sp = s_sub[t] - b;
    for (int i = lo; i < hi; ++i) {
        const int c = counts[i];
        offsets[i] = pos;
        for (int s = 0; s * maxpts < c; ++s) sub[sp++] = make_int4(i, pos + s * maxpts, min(maxpts, c - s * maxpts), 0);
        pos += c;
    }
    if (t == kScanThreads - 1) *nsub = s_sub[t];
}

template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const T* __restrict__ x, const T* __restrict__ y, int M, int L1,
                                                 int L2, int TS, int tiles1, const int* __restrict__ rank,
                                                 const int* __restrict__ offsets, cplx<T>* __restrict__ xy,
                                                 int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const T xi = x[i], yi = y[i];
    int m1, m2;
    T t1, t2;
    locate(xi, L1, m1, t1);
    locate(yi, L2, m2, t2);
    const int bin = (m2 / TS) * tiles1 + m1 / TS;
    const int p = offsets[bin] + rank[i];
    xy[p] = mk(xi, yi);
    idx[p] = i;
}

// ================================================================= gather / spread
// Window weights for one point: w1 along dim 1, w2 along dim 2.
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

template <typename T> constexpr int tile_size() { return sizeof(T) == 4 ? 32 : 16; }

// Load the (TS+2K)^2 region of the transposed grid around tile (t1, t2) into LDS.
template <typename C, int TS>
__device__ __forceinline__ void load_tile(C* tile, const C* __restrict__ Gt, int L1, int L2, int o1, int o2)
{
    constexpr int W = TS + 2 * kK;
    for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
        const int r = e / W, c = e - r * W;
        tile[e] = Gt[(size_t)wrap(o2 + r, L2) * L1 + wrap(o1 + c, L1)];
    }
}

// (c) type-2 gather, MODE 0: out[l] = y_l;  MODE 1: out = y, out_d1 = dy/dx1,
// out_d2 = dy/dx2;  MODE 2: gradient Re(conj(g_l) dy_l/dx) into gx1/gx2.
// y_l = sum_{m2} w2[m2] sum_{m1} w1[m1] G[(mu2+m2-K) mod L2][(mu1+m1-K) mod L1]
// (BP_FBP_FFBP/NFFT1_2D.m:61-67 on the unshifted grid).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_gather(
    const cplx<T>* __restrict__ Gt, int L1, int L2, const cplx<T>* __restrict__ xy,
    const int* __restrict__ idx, const int4* __restrict__ sub, const int* __restrict__ nsub, int tiles1,
    WinArgs wa, cplx<T>* __restrict__ out, cplx<T>* __restrict__ out_d1, cplx<T>* __restrict__ out_d2,
    const cplx<T>* __restrict__ g, T* __restrict__ gx1, T* __restrict__ gx2)
{
    using C = cplx<T>;
    constexpr int TS = tile_size<T>();
    constexpr int W = TS + 2 * kK;
    __shared__ C tile[W * W];
    const int s = blockIdx.x;
    if (s >= *nsub) return;
    const int4 d = sub[s];
    const int t1 = d.x % tiles1, t2 = d.x / tiles1;
    load_tile<C, TS>(tile, Gt, L1, L2, t1 * TS - kK, t2 * TS - kK);
    __syncthreads();
    for (int p = d.y + threadIdx.x; p < d.y + d.z; p += blockDim.x) {
        const C q = xy[p];
        int m1, m2;
        T f1, f2;
        locate(q.x, L1, m1, f1);
        locate(q.y, L2, m2, f2);
        const int c0 = min(max(m1 - t1 * TS, 0), TS - 1);
        const int r0 = min(max(m2 - t2 * TS, 0), TS - 1);
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
            out[idx[p]] = mk(ax, ay);
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
            const int o = idx[p];
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
}

// (d) type-1 spread: each block owns one subproblem (<= maxpts points of one
// tile), accumulates the tile + K-halo in LDS with LDS atomics, then adds it
// into the grid (one global atomic per touched cell instead of the
// reference's 2 (2K+1)^2 per point, NFFT/NFFT2_2D.cu:127-128).
template <typename T>
__global__ __launch_bounds__(256) void k_spread(
    cplx<T>* __restrict__ Gt, int L1, int L2, const cplx<T>* __restrict__ xy, const int* __restrict__ idx,
    const cplx<T>* __restrict__ data, const int4* __restrict__ sub, const int* __restrict__ nsub, int tiles1,
    WinArgs wa)
{
    using C = cplx<T>;
    constexpr int TS = tile_size<T>();
    constexpr int W = TS + 2 * kK;
    __shared__ T tre[W * W];
    __shared__ T tim[W * W];
    const int s = blockIdx.x;
    if (s >= *nsub) return;
    const int4 d = sub[s];
    const int t1 = d.x % tiles1, t2 = d.x / tiles1;
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
        locate(q.x, L1, m1, f1);
        locate(q.y, L2, m2, f2);
        const int c0 = min(max(m1 - t1 * TS, 0), TS - 1);
        const int r0 = min(max(m2 - t2 * TS, 0), TS - 1);
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
    const int o1 = t1 * TS - kK, o2 = t2 * TS - kK;
    for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
        const T a = tre[e], b = tim[e];
        if (a != (T)0 || b != (T)0) {
            const int r = e / W, c = e - r * W;
            C* cell = &Gt[(size_t)wrap(o2 + r, L2) * L1 + wrap(o1 + c, L1)];
            unsafeAtomicAdd(&cell->x, a);
            unsafeAtomicAdd(&cell->y, b);
        }
    }
}

}  // namespace nfft2d
