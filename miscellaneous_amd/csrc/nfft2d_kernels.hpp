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

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2.  Give every XCD a contiguous range of logical
// blocks so that neighbouring column blocks, which share cache lines, meet in
// the same L2.  Identity when the grid is not a multiple of 8.
__device__ __forceinline__ int xcd_block(int b, int n)
{
    if (n & 7) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

// For work of uneven cost in logical order (subproblems of a point set's
// bins): every XCD takes runs of 8 consecutive logical blocks, the runs dealt
// round-robin, so that neighbouring subproblems share an L2 while the XCDs
// stay evenly loaded.  Grid: a multiple of 64.
__device__ __forceinline__ int xcd_runs(int b)
{
    return (b & ~63) | ((b & 7) << 3) | ((b >> 3) & 7);
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
template <typename T, int N> constexpr int t2_rows_lds() { return rows_per_block<N>() * 2 * lds_line<N>(); }
template <typename T, int N, int SIGN>
__device__ __forceinline__ void t2_rows_body(int blk, cplx<T>* lds, const cplx<T>* __restrict__ f,
                                             cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
                                             const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    using C = cplx<T>;
    constexpr int ROWS = rows_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    const int g = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int i1 = blk * ROWS + g;
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
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(rows_per_block<N>() * (N / 8)) void k_t2_rows(
    const cplx<T>* __restrict__ f, cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    __shared__ cplx<T> lds[t2_rows_lds<T, N>()];
    t2_rows_body<T, N, SIGN>(blockIdx.x, lds, f, Tm, D1, D2, tw, N1);
}

// (b) column pass: Gt[n2][n1] = sum_{k1} T[k1][n2] e^{-2 pi i k1 n1 / L1}, n1 in [0, L1).
// Loads B columns of T per block (B consecutive elements per row), writes B
// contiguous rows of the transposed grid.
template <typename T, int N> constexpr int t2_cols_lds() { return cols_per_block<N>() * 2 * lds_line<N>(); }
template <typename T, int N, int SIGN>
__device__ __forceinline__ void t2_cols_body(int blk, cplx<T>* lds, const cplx<T>* __restrict__ Tm,
                                             cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int L2)
{
    using C = cplx<T>;
    constexpr int B = cols_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    const int t = threadIdx.x;
    const int bF = t % B, jF = t / B;      // first stage: B columns side by side
    const int bL = t / TPL, jL = t % TPL;  // last stage: whole lines
    const int n2F = min(blk * B + bF, L2 - 1);
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
    const int n2L = blk * B + bL;
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
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(cols_per_block<N>() * (N / 8)) void k_t2_cols(
    const cplx<T>* __restrict__ Tm, cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int L2)
{
    __shared__ cplx<T> lds[t2_cols_lds<T, N>()];
    t2_cols_body<T, N, SIGN>(xcd_block(blockIdx.x, gridDim.x), lds, Tm, Gt, tw, L2);
}

// ================================================================= type-1 passes
// Column pass: Tp[k1][n2] = sum_{n1} Gt[n2][n1] e^{SIGN 2 pi i n1 k1 / L1} for
// k1 in [-N1/2, N1/2) only (the reference crops the rest, NFFT/NFFT2_2D.cu:74-86).
template <typename T, int N> constexpr int t1_cols_lds() { return cols_per_block<N>() * 2 * lds_line<N>(); }
template <typename T, int N, int SIGN>
__device__ __forceinline__ void t1_cols_body(int blk, cplx<T>* lds, const cplx<T>* __restrict__ Gt,
                                             cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int L2)
{
    using C = cplx<T>;
    constexpr int B = cols_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    const int t = threadIdx.x;
    const int bF = t / TPL, jF = t % TPL;  // first stage: whole grid rows (contiguous)
    const int bL = t % B, jL = t / B;      // last stage: B outputs side by side
    const int n2F = min(blk * B + bF, L2 - 1);
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
    const int n2L = blk * B + bL;
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
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(cols_per_block<N>() * (N / 8)) void k_t1_cols(
    const cplx<T>* __restrict__ Gt, cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int L2)
{
    __shared__ cplx<T> lds[t1_cols_lds<T, N>()];
    t1_cols_body<T, N, SIGN>(xcd_block(blockIdx.x, gridDim.x), lds, Gt, Tp, tw, L2);
}

// Row pass + crop + deconvolution: F[k1][k2] = D1[k1] D2[k2] sum_{n2} Tp[k1][n2] e^{SIGN 2 pi i n2 k2 / L2}
// (NFFT/NFFT2_2D.cu:74-100 Decimation_and_Scaling fused into the last FFT stage).
template <typename T, int N> constexpr int t1_rows_lds() { return rows_per_block<N>() * 2 * lds_line<N>(); }
template <typename T, int N, int SIGN>
__device__ __forceinline__ void t1_rows_body(int blk, cplx<T>* lds, const cplx<T>* __restrict__ Tp,
                                             cplx<T>* __restrict__ F, const T* __restrict__ D1,
                                             const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    using C = cplx<T>;
    constexpr int ROWS = rows_per_block<N>();
    constexpr int TPL = N / 8, LL = lds_line<N>(), V = 2;
    constexpr int R0 = FftPlan<N>::radix(0);
    constexpr int RL = FftPlan<N>::radix(FftPlan<N>::stages - 1);
    const int g = threadIdx.x / TPL, j = threadIdx.x % TPL;
    const int i1 = blk * ROWS + g;
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
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(rows_per_block<N>() * (N / 8)) void k_t1_rows(
    const cplx<T>* __restrict__ Tp, cplx<T>* __restrict__ F, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1)
{
    __shared__ cplx<T> lds[t1_rows_lds<T, N>()];
    t1_rows_body<T, N, SIGN>(blockIdx.x, lds, Tp, F, D1, D2, tw, N1);
}

// ================================================================= direct-DFT passes
// Twiddle table of a direct pass: staged into LDS when the caller has room
// (lds_tw non-null), else read from global memory.
template <typename T>
__device__ __forceinline__ const cplx<T>* stage_twiddles(const cplx<T>* __restrict__ tw, cplx<T>* lds_tw, int L)
{
    if (!lds_tw) return tw;
    for (int k = threadIdx.x; k < L; k += blockDim.x) lds_tw[k] = tw[k];
    return lds_tw;
}

// sum_{p < P, p = r mod G} x[p] w[(e0 + p step) mod L] for this thread's r,
// in four interleaved partial sums (independent FMA chains), added as
// (s0 + s1) + (s2 + s3).
template <int SIGN, typename T>
__device__ __forceinline__ cplx<T> direct_sum(const cplx<T>* x, int P, const cplx<T>* w, int e0, int step, int L,
                                              int r, int G)
{
    using C = cplx<T>;
    C s0 = mk((T)0, (T)0), s1 = s0, s2 = s0, s3 = s0;
    int e = (int)((e0 + (long long)r * step) % L);
    const int st = (int)(((long long)G * step) % L);
    auto adv = [&]() {
        e += st;
        e -= e >= L ? L : 0;
    };
    int p = r;
    for (; p + 3 * G < P; p += 4 * G) {
        s0 = cadd(s0, cmul(x[p], twid<SIGN>(w, e)));
        adv();
        s1 = cadd(s1, cmul(x[p + G], twid<SIGN>(w, e)));
        adv();
        s2 = cadd(s2, cmul(x[p + 2 * G], twid<SIGN>(w, e)));
        adv();
        s3 = cadd(s3, cmul(x[p + 3 * G], twid<SIGN>(w, e)));
        adv();
    }
    for (; p < P; p += G) {
        s0 = cadd(s0, cmul(x[p], twid<SIGN>(w, e)));
        adv();
    }
    return cadd(cadd(s0, s1), cadd(s2, s3));
}

// All Nout outputs of a direct pass over the block: for short lines G = 2 or
// 4 adjacent threads share an output (each sums every G-th term, then they
// add their partials by cross-lane shuffles), so that the block's lanes stay
// busy.  G depends on Nout only, never on the block size, so that every
// launch shape (plans: 256 threads, batch row passes: 128) gives the same bits.
// term(o, e0, step) gives output o's twiddle start and step; store(o, v).
template <int SIGN, typename T, typename Term, typename Store>
__device__ __forceinline__ void direct_outputs(const cplx<T>* x, int P, const cplx<T>* w, int L, int Nout,
                                               Term&& term, Store&& store)
{
    const int NT = blockDim.x;
    const int G = Nout <= 64 ? 4 : (Nout <= 128 ? 2 : 1);   // NT >= 128 >= G
    const int r = (int)threadIdx.x % G;
    for (int o = (int)threadIdx.x / G; o < Nout; o += NT / G) {
        int e0, step;
        term(o, e0, step);
        cplx<T> v = direct_sum<SIGN, T>(x, P, w, e0, step, L, r, G);
        for (int m = 1; m < G; m <<= 1) {
            v.x += __shfl_xor(v.x, m, 64);
            v.y += __shfl_xor(v.y, m, 64);
        }
        if (r == 0) store(o, v);
    }
}

// Direct DFT passes for line lengths the Stockham passes do not take: each
// output a sum over the line staged in LDS, with twiddle index (k n) mod L
// into the 2N-entry table advanced term by term.  O(N^2) per line; used for
// the reference's non-power-of-two sizes below the chirp-z threshold (e.g.
// 2Nu = 88 in Synthesis/PSO).  lds_tw: room for the L twiddles in LDS, or null.
template <typename T, int SIGN>
__device__ __forceinline__ void t2_rows_direct_body(int i1, cplx<T>* line, const cplx<T>* __restrict__ f,
                                                    cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
                                                    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1,
                                                    int N2, cplx<T>* lds_tw = nullptr)
{
    const int L2 = 2 * N2;
    const cplx<T>* w = stage_twiddles<T>(tw, lds_tw, L2);
    for (int i2 = threadIdx.x; i2 < N2; i2 += blockDim.x)
        line[i2] = cscale(f[(size_t)i1 * N2 + i2], D1[i1] * D2[i2]);
    __syncthreads();
    // twiddle index ((k2 mod L2) n2) mod L2 for k2 = i2 - N2/2: starts at (L2 - N2/2) n2, steps by n2
    direct_outputs<SIGN, T>(
        line, N2, w, L2, L2,
        [&](int n2, int& e0, int& step) {
            e0 = (int)(((long long)(L2 - N2 / 2) * n2) % L2);
            step = n2;
        },
        [&](int n2, cplx<T> v) { Tm[(size_t)i1 * L2 + n2] = v; });
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t2_rows_direct(
    const cplx<T>* __restrict__ f, cplx<T>* __restrict__ Tm, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1, int N2)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    t2_rows_direct_body<T, SIGN>(blockIdx.x, lds, f, Tm, D1, D2, tw, N1, N2, lds + N2);
}

template <typename T, int SIGN>
__device__ __forceinline__ void t2_cols_direct_body(int n2, cplx<T>* line, const cplx<T>* __restrict__ Tm,
                                                    cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int N1,
                                                    int L2, cplx<T>* lds_tw = nullptr)
{
    const int L1 = 2 * N1;
    const cplx<T>* w = stage_twiddles<T>(tw, lds_tw, L1);
    for (int i1 = threadIdx.x; i1 < N1; i1 += blockDim.x) line[i1] = Tm[(size_t)i1 * L2 + n2];
    __syncthreads();
    direct_outputs<SIGN, T>(
        line, N1, w, L1, L1,
        [&](int n1, int& e0, int& step) {
            e0 = (int)(((long long)(L1 - N1 / 2) * n1) % L1);
            step = n1;
        },
        [&](int n1, cplx<T> v) { Gt[(size_t)n2 * L1 + n1] = v; });
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t2_cols_direct(
    const cplx<T>* __restrict__ Tm, cplx<T>* __restrict__ Gt, const cplx<T>* __restrict__ tw, int N1, int L2)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    t2_cols_direct_body<T, SIGN>(blockIdx.x, lds, Tm, Gt, tw, N1, L2, lds + N1);
}

template <typename T, int SIGN>
__device__ __forceinline__ void t1_cols_direct_body(int n2, cplx<T>* line, const cplx<T>* __restrict__ Gt,
                                                    cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int N1,
                                                    int L2, cplx<T>* lds_tw = nullptr)
{
    const int L1 = 2 * N1;
    const cplx<T>* w = stage_twiddles<T>(tw, lds_tw, L1);
    for (int n1 = threadIdx.x; n1 < L1; n1 += blockDim.x) line[n1] = Gt[(size_t)n2 * L1 + n1];
    __syncthreads();
    direct_outputs<SIGN, T>(
        line, L1, w, L1, N1,
        [&](int i1, int& e0, int& step) {
            const int k1 = i1 - N1 / 2;
            e0 = 0;
            step = k1 < 0 ? k1 + L1 : k1;
        },
        [&](int i1, cplx<T> v) { Tp[(size_t)i1 * L2 + n2] = v; });
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t1_cols_direct(
    const cplx<T>* __restrict__ Gt, cplx<T>* __restrict__ Tp, const cplx<T>* __restrict__ tw, int N1, int L2)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    t1_cols_direct_body<T, SIGN>(blockIdx.x, lds, Gt, Tp, tw, N1, L2, lds + 2 * N1);
}

template <typename T, int SIGN>
__device__ __forceinline__ void t1_rows_direct_body(int i1, cplx<T>* line, const cplx<T>* __restrict__ Tp,
                                                    cplx<T>* __restrict__ F, const T* __restrict__ D1,
                                                    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1,
                                                    int N2, cplx<T>* lds_tw = nullptr)
{
    const int L2 = 2 * N2;
    const cplx<T>* w = stage_twiddles<T>(tw, lds_tw, L2);
    for (int n2 = threadIdx.x; n2 < L2; n2 += blockDim.x) line[n2] = Tp[(size_t)i1 * L2 + n2];
    __syncthreads();
    direct_outputs<SIGN, T>(
        line, L2, w, L2, N2,
        [&](int i2, int& e0, int& step) {
            const int k2 = i2 - N2 / 2;
            e0 = 0;
            step = k2 < 0 ? k2 + L2 : k2;
        },
        [&](int i2, cplx<T> v) { F[(size_t)i1 * N2 + i2] = cscale(v, D1[i1] * D2[i2]); });
}

template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_t1_rows_direct(
    const cplx<T>* __restrict__ Tp, cplx<T>* __restrict__ F, const T* __restrict__ D1,
    const T* __restrict__ D2, const cplx<T>* __restrict__ tw, int N1, int N2)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    t1_rows_direct_body<T, SIGN>(blockIdx.x, lds, Tp, F, D1, D2, tw, N1, N2, lds + 2 * N2);
}

// ================================================================= type-1 direct passes on MFMA
// The type-1 passes of short complex64 lines that are not powers of two (up
// to kT1MmMaxN modes, below the chirp-z threshold: e.g. 88 in Synthesis/PSO)
// as complex matrix products on the FP32 matrix cores (v_mfma_f32_16x16x4f32:
// A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k, accumulator element
// r of lane l at row 4 (l / 16) + r, column l % 16).  Column pass: Tm (N1 x
// L2) = W1 (N1 x L1) . G^T, W1[i1][n1] = w((k1 n1) mod L1), k1 = i1 - N1/2;
// row pass: F (N1 x N2) = D1 D2 (Tm (N1 x L2) . W2^T), W2[i2][n2] = w((k2 n2)
// mod L2).  A block takes 16 lines (staged in LDS, rows padded to an odd
// length) and their 16 x 16 output tiles; the twiddle operand is read from
// the LDS twiddle table at an index advanced by 4 k per K-step of 4, and four
// real products (Re Re, Im Im, Re Im, Im Re) accumulate per K-step.  The VALU
// direct passes read a line value and a twiddle from LDS for every 4 FMAs.
constexpr int kT1MmMaxN = 126;
typedef float mm4v __attribute__((ext_vector_type(4)));
constexpr int kT1MmLd = 2 * kT1MmMaxN + 1;                 // padded line of the LDS panel (odd)
constexpr int kT1MmLds = 16 * kT1MmLd + 2 * kT1MmMaxN;     // panel + twiddles, float2 entries

// 16 x 16 complex tile, K-steps of 4 over K = L (a multiple of 4), this wave
// taking k0 = kf, kf + ks, ...: lane (i, kk) = (l % 16, l / 16).  Left operand
// a(k0) and right operand b(k0) of this lane.
template <typename AF, typename BF>
__device__ __forceinline__ void mm_tile(int L, int kf, int ks, AF&& a_of, BF&& b_of, mm4v& rr, mm4v& ii, mm4v& ri,
                                        mm4v& ir)
{
    for (int k0 = kf; k0 < L; k0 += ks) {
        const float2 a = a_of(k0), b = b_of(k0);
        rr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, rr, 0, 0, 0);
        ii = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, ii, 0, 0, 0);
        ri = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, ri, 0, 0, 0);
        ir = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, ir, 0, 0, 0);
    }
}

// One output tile over this wave's K-steps.  The twiddle line of lane i is
// that of frequency index 16 t + i (k = that - N/2) over L = 2N samples; the
// panel holds the 16 data lines.  Twiddles are the left operand of the column
// pass (Tm = W1 . G^T) and the right operand of the row pass (F = Tm . W2^T).
template <int SIGN, bool TW_LEFT>
__device__ __forceinline__ void t1_mm_tile(int t, const float2* panel, const float2* wt, int N, int kf, int ks,
                                           mm4v& rr, mm4v& ii, mm4v& ri, mm4v& ir)
{
    const int L = 2 * N;
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
    const int f = 16 * t + i;
    const int k = f - N / 2, step = k < 0 ? k + L : k;
    const int adv = (step * ks) % L;
    int e = (step * (kf + kk)) % L;                          // twiddle index (k n) mod L, n = k0 + kk
    const bool in = f < N;
    auto tw = [&](int) {
        const float2 w = in ? twid<SIGN>(wt, e) : make_float2(0.f, 0.f);
        e += adv;
        e -= e >= L ? L : 0;
        return w;
    };
    auto line = [&](int k0) { return panel[i * kT1MmLd + k0 + kk]; };
    rr = ii = ri = ir = mm4v{0.f, 0.f, 0.f, 0.f};
    if constexpr (TW_LEFT)
        mm_tile(L, kf, ks, tw, line, rr, ii, ri, ir);
    else
        mm_tile(L, kf, ks, line, tw, rr, ii, ri, ir);
}

// Stores of a column-pass tile (row tile rt of i1, column tile ct of n2) and
// of a row-pass tile (row tile rt of i1, column tile ct of i2, deconvolved).
__device__ __forceinline__ void t1_cols_mm_store(int rt, int ct, const mm4v& rr, const mm4v& ii, const mm4v& ri,
                                                 const mm4v& ir, float2* __restrict__ Tm, int N1, int L2)
{
    const int lane = threadIdx.x & 63, n2 = 16 * ct + (lane & 15), kk = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int io = 16 * rt + 4 * kk + r;
        if (io < N1 && n2 < L2) Tm[(size_t)io * L2 + n2] = make_float2(rr[r] - ii[r], ri[r] + ir[r]);
    }
}
__device__ __forceinline__ void t1_rows_mm_store(int rt, int ct, const mm4v& rr, const mm4v& ii, const mm4v& ri,
                                                 const mm4v& ir, float2* __restrict__ F, const float* __restrict__ D1,
                                                 const float* __restrict__ D2, int N1, int N2)
{
    const int lane = threadIdx.x & 63, i2 = 16 * ct + (lane & 15), kk = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i1 = 16 * rt + 4 * kk + r;
        if (i1 < N1 && i2 < N2)
            F[(size_t)i1 * N2 + i2] = cscale(make_float2(rr[r] - ii[r], ri[r] + ir[r]), D1[i1] * D2[i2]);
    }
}

// Column pass of column tile ct, the 16 lines (rows n2 = 16 ct + j of the
// grid) in panel[j * kT1MmLd + n1] and the L1 twiddles in wt: Tm[i1][n2] for
// every row tile of i1, one tile per wave at a time (w, w + waves, ...).
template <int SIGN>
__device__ __forceinline__ void t1_cols_mm_compute(int ct, const float2* panel, const float2* wt,
                                                   float2* __restrict__ Tm, int N1, int L2)
{
    for (int rt = threadIdx.x >> 6; rt < (N1 + 15) / 16; rt += blockDim.x >> 6) {
        mm4v rr, ii, ri, ir;
        t1_mm_tile<SIGN, true>(rt, panel, wt, N1, 0, 4, rr, ii, ri, ir);
        t1_cols_mm_store(rt, ct, rr, ii, ri, ir, Tm, N1, L2);
    }
}

// Row pass of row tile rt, the 16 lines (rows i1 = 16 rt + j of Tm) in
// panel[j * kT1MmLd + n2] and the L2 twiddles in wt: F[i1][i2], deconvolved,
// for every column tile of i2, one tile per wave at a time.
template <int SIGN>
__device__ __forceinline__ void t1_rows_mm_compute(int rt, const float2* panel, const float2* wt,
                                                   float2* __restrict__ F, const float* __restrict__ D1,
                                                   const float* __restrict__ D2, int N1, int N2)
{
    for (int ct = threadIdx.x >> 6; ct < (N2 + 15) / 16; ct += blockDim.x >> 6) {
        mm4v rr, ii, ri, ir;
        t1_mm_tile<SIGN, false>(ct, panel, wt, N2, 0, 4, rr, ii, ri, ir);
        t1_rows_mm_store(rt, ct, rr, ii, ri, ir, F, D1, D2, N1, N2);
    }
}

// 16 rows of a row-major matrix (row stride len <= 2 kT1MmMaxN, nrows rows)
// from row r0 on, zero past the last, into the panel; and the len twiddles
// into wt.  Blocks of 256 threads; every load is issued before the first LDS
// store (a loop of dependent load-store rounds took 8.3 us a pass at 88 x 88
// against 6.5 us for the VALU direct pass).
constexpr int kT1MmStage = (16 * 2 * kT1MmMaxN + 255) / 256;
__device__ __forceinline__ void mm_stage_rows(const float2* __restrict__ src, int r0, int nrows, int len,
                                              const float2* __restrict__ tw, float2* panel, float2* wt)
{
    float2 v[kT1MmStage];
#pragma unroll
    for (int q = 0; q < kT1MmStage; ++q) {
        const int e = threadIdx.x + 256 * q, j = e / len, c = e - j * len, r = r0 + j;
        v[q] = e < 16 * len && r < nrows ? src[(size_t)r * len + c] : make_float2(0.f, 0.f);
    }
    const float2 w = threadIdx.x < len ? tw[threadIdx.x] : make_float2(0.f, 0.f);
#pragma unroll
    for (int q = 0; q < kT1MmStage; ++q) {
        const int e = threadIdx.x + 256 * q, j = e / len, c = e - j * len;
        if (e < 16 * len) panel[j * kT1MmLd + c] = v[q];
    }
    if (threadIdx.x < len) wt[threadIdx.x] = w;
}

// The 4 waves' partial tiles summed in LDS (wave 0 adds waves 1, 2, 3 in turn).
__device__ __forceinline__ bool mm_sum_waves(mm4v* part, mm4v& rr, mm4v& ii, mm4v& ri, mm4v& ir)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (w > 0) {
        mm4v* q = part + (w - 1) * 256 + lane;
        q[0] = rr;
        q[64] = ii;
        q[128] = ri;
        q[192] = ir;
    }
    __syncthreads();
    if (w > 0) return false;
    for (int v = 0; v < 3; ++v) {
        const mm4v* q = part + v * 256 + lane;
        rr += q[0];
        ii += q[64];
        ri += q[128];
        ir += q[192];
    }
    return true;
}

// A plan's passes: one block per 16 x 16 output tile (a single transform has
// too few tiles to fill the chip one tile per wave), its 4 waves taking the
// K-steps w, w + 4, ... and their partial tiles summed in LDS.  Column pass:
// blockIdx.x the column tile of n2, blockIdx.y the row tile of i1; row pass:
// blockIdx.x the row tile of i1, blockIdx.y the column tile of i2.
template <int SIGN>
__global__ __launch_bounds__(256) void k_t1_cols_mm(const float2* __restrict__ Gt, float2* __restrict__ Tm,
                                                    const float2* __restrict__ tw1, int N1, int L2)
{
    __shared__ float2 lds[kT1MmLds];
    __shared__ mm4v part[3 * 256];
    mm_stage_rows(Gt, 16 * blockIdx.x, L2, 2 * N1, tw1, lds, lds + 16 * kT1MmLd);
    __syncthreads();
    mm4v rr, ii, ri, ir;
    t1_mm_tile<SIGN, true>(blockIdx.y, lds, lds + 16 * kT1MmLd, N1, 4 * (threadIdx.x >> 6), 16, rr, ii, ri, ir);
    if (mm_sum_waves(part, rr, ii, ri, ir)) t1_cols_mm_store(blockIdx.y, blockIdx.x, rr, ii, ri, ir, Tm, N1, L2);
}
template <int SIGN>
__global__ __launch_bounds__(256) void k_t1_rows_mm(const float2* __restrict__ Tm, float2* __restrict__ F,
                                                    const float* __restrict__ D1, const float* __restrict__ D2,
                                                    const float2* __restrict__ tw2, int N1, int N2)
{
    __shared__ float2 lds[kT1MmLds];
    __shared__ mm4v part[3 * 256];
    mm_stage_rows(Tm, 16 * blockIdx.x, N1, 2 * N2, tw2, lds, lds + 16 * kT1MmLd);
    __syncthreads();
    mm4v rr, ii, ri, ir;
    t1_mm_tile<SIGN, false>(blockIdx.y, lds, lds + 16 * kT1MmLd, N2, 4 * (threadIdx.x >> 6), 16, rr, ii, ri, ir);
    if (mm_sum_waves(part, rr, ii, ri, ir))
        t1_rows_mm_store(blockIdx.x, blockIdx.y, rr, ii, ri, ir, F, D1, D2, N1, N2);
}

// ================================================================= setpts (bin sort)
// Points are counting-sorted by bin: a bin is a kBin x kBin block of cells of
// the oversampled grid, holding mu = rint(c x) mod L.  Bins are the tiles of the
// gather and the output tiles of the spread.
constexpr int kBin = 16;
constexpr int kMaxLdsBins = 16384;      // per-block LDS histogram limit (64 KiB)

// A sorted point: fractional offsets t = c x - rint(c x) (exact), the packed
// cell mu2 << 16 | mu1 and the point's original index.  One 16-byte (FP32) or
// 32-byte (FP64) load per point.
template <typename T> struct PtRec;
template <> struct __attribute__((aligned(16))) PtRec<float> {
    float t1, t2;
    int cell, idx;
};
template <> struct __attribute__((aligned(16))) PtRec<double> {
    double t1, t2;
    int cell, idx, pad0, pad1;
};

template <typename T>
__device__ __forceinline__ int bin_of(T x, T y, int L1, int L2, T inv1, T inv2, int nb1, int& m1, int& m2, T& t1,
                                      T& t2)
{
    locate(x, L1, inv1, m1, t1);
    locate(y, L2, inv2, m2, t2);
    return (m2 / kBin) * nb1 + m1 / kBin;
}

// Rank of this lane's point among the wave's points of the same bin, with one
// counter increment per run of equal bins in consecutive lanes (points that
// arrive in coordinate order, as along the rays of a polar grid, share bins
// over long runs).  Inactive lanes pass act = false.  Returns the point's rank
// in counter[bin].
__device__ __forceinline__ int wave_run_rank(int* counter, int bin, bool act)
{
    const int lane = threadIdx.x & 63;
    const int key = act ? bin : -1 - lane;                // never equal to a neighbour
    const int prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || key != prev;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int s = 63 - __clzll(heads & upto);             // head of this lane's run
    int base = 0;
    if (head && act) {
        const unsigned long long rest = lane == 63 ? 0ull : (heads >> (lane + 1));
        const int next = rest ? lane + __ffsll((long long)rest) : 64;
        base = atomicAdd(&counter[bin], next - lane);
    }
    base = __shfl(base, s, 64);
    return base + (lane - s);
}

// Pass 1: block b counts the bins of points [b*chunk, (b+1)*chunk) in an LDS
// histogram (integer LDS atomics, one per run of equal bins in a wave);
// rank[i] is the point's rank inside its (block, bin), written unless rank is
// null (k_scatter_lds recomputes it).  hist is [nblk][nbins].
constexpr int kHistThreads = 1024;   // 16 waves per block: 4 per SIMD to hide the load latency
template <typename T>
__device__ __forceinline__ void hist_body(int blk, int* h, const T* __restrict__ x, const T* __restrict__ y, int M,
                                          int chunk, int L1, int L2, T inv1, T inv2, int nb1, int nbins,
                                          int* __restrict__ hist, int* __restrict__ rank)
{
    for (int i = threadIdx.x; i < nbins; i += kHistThreads) h[i] = 0;
    __syncthreads();
    // batches of kHistBatch points per thread, all loads of a batch in flight
    // together (a loop with one load per iteration pays the full memory
    // latency every time)
    constexpr int kHistBatch = 4;
    const int lo = blk * chunk, hi = min(lo + chunk, M);
    for (int b = lo; b < hi; b += kHistBatch * kHistThreads) {
        T xv[kHistBatch], yv[kHistBatch];
#pragma unroll
        for (int q = 0; q < kHistBatch; ++q) {
            const int i = b + q * kHistThreads + (int)threadIdx.x;
            xv[q] = i < hi ? x[i] : (T)0;
            yv[q] = i < hi ? y[i] : (T)0;
        }
#pragma unroll
        for (int q = 0; q < kHistBatch; ++q) {
            const int i = b + q * kHistThreads + (int)threadIdx.x;
            const bool act = i < hi;
            int m1, m2;
            T t1, t2;
            const int bin = bin_of(xv[q], yv[q], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
            const int r = wave_run_rank(h, bin, act);
            if (act && rank != nullptr) rank[i] = r;
        }
    }
    __syncthreads();
    int* out = hist + (size_t)blk * nbins;
    for (int i = threadIdx.x; i < nbins; i += kHistThreads) out[i] = h[i];
}
template <typename T>
__global__ __launch_bounds__(kHistThreads) void k_hist(const T* __restrict__ x, const T* __restrict__ y, int M, int chunk,
                                                       int L1, int L2, T inv1, T inv2, int nb1, int nbins,
                                                       int* __restrict__ hist, int* __restrict__ rank)
{
    extern __shared__ int h[];
    hist_body<T>(blockIdx.x, h, x, y, M, chunk, L1, L2, inv1, inv2, nb1, nbins, hist, rank);
}

// Pass 1 for grids with more than kMaxLdsBins bins: one logical block, global
// integer atomics (one per run of equal bins in a wave).
template <typename T>
__global__ __launch_bounds__(256) void k_hist_global(const T* __restrict__ x, const T* __restrict__ y, int M, int L1,
                                                     int L2, T inv1, T inv2, int nb1, int* __restrict__ hist,
                                                     int* __restrict__ rank)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < M;
    int bin = 0;
    if (act) {
        int m1, m2;
        T t1, t2;
        bin = bin_of(x[i], y[i], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
    }
    const int r = wave_run_rank(hist, bin, act);
    if (act) rank[i] = r;
}

// Pass 2: per bin, turn hist[blk][bin] into the offset of (blk, bin) inside
// the bin and write the bin total.  A block covers 32 bins x 8 slices of the
// block range; each thread's loads are independent and coalesced across bins.
// The block also writes, for its group of 32 bins, the exclusive prefix of
// points and subproblems inside the group (gexcl, as int2) and the group sums
// (gsum[blockIdx.x], as int2).
constexpr int kColBins = 32, kColSlices = 8;   // 32 bins: 128-byte rows of hist
constexpr int kColPer = 256 / kColSlices;         // block-range slice held in registers (nblk <= 256)
__device__ __forceinline__ void colscan_body(int blk, int* __restrict__ hist, int nblk, int nbins, int maxpts,
                                             int* __restrict__ tot, int2* __restrict__ gexcl, int2* __restrict__ gsum)
{
    __shared__ int part[kColSlices][kColBins + 1];
    const int bl = threadIdx.x % kColBins, q = threadIdx.x / kColBins;
    const int bin = blk * kColBins + bl;
    const int per = (nblk + kColSlices - 1) / kColSlices;
    const int lo = min(q * per, nblk), hi = min(lo + per, nblk);
    int v[kColPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kColPer; ++k) {
        const int kb = lo + k;
        v[k] = (bin < nbins && kb < hi) ? hist[(size_t)kb * nbins + bin] : 0;
        sum += v[k];
    }
    for (int kb = lo + kColPer; kb < hi; ++kb) sum += bin < nbins ? hist[(size_t)kb * nbins + bin] : 0;
    part[q][bl] = sum;
    __syncthreads();
    if (q == 0) {
        int run = 0;
        for (int j = 0; j < kColSlices; ++j) {
            const int t = part[j][bl];
            part[j][bl] = run;
            run += t;
        }
        if (bin < nbins) tot[bin] = run;
        // exclusive prefix over the 32 bins of the group (lanes 0..31 of wave 0)
        const int ns = (run + maxpts - 1) / maxpts;
        int ia = run, ib = ns;
#pragma unroll
        for (int o = 1; o < kColBins; o <<= 1) {
            const int na = __shfl_up(ia, o, kColBins), nb = __shfl_up(ib, o, kColBins);
            if (bl >= o) { ia += na; ib += nb; }
        }
        if (bin < nbins) gexcl[bin] = make_int2(ia - run, ib - ns);
        if (bl == kColBins - 1) gsum[blk] = make_int2(ia, ib);
    }
    __syncthreads();
    if (bin < nbins) {
        int run = part[q][bl];
#pragma unroll
        for (int k = 0; k < kColPer; ++k) {
            const int kb = lo + k;
            if (kb < hi) {
                hist[(size_t)kb * nbins + bin] = run;
                run += v[k];
            }
        }
        for (int kb = lo + kColPer; kb < hi; ++kb) {
            const int t = hist[(size_t)kb * nbins + bin];
            hist[(size_t)kb * nbins + bin] = run;
            run += t;
        }
    }
}

__global__ __launch_bounds__(256) void k_colscan(int* __restrict__ hist, int nblk, int nbins, int maxpts,
                                                 int* __restrict__ tot, int2* __restrict__ gexcl,
                                                 int2* __restrict__ gsum)
{
    colscan_body(blockIdx.x, hist, nblk, nbins, maxpts, tot, gexcl, gsum);
}

// Pass 2b (one block): exclusive scan of the group sums in place, in rounds of
// kScanThreads groups; the totals go to nsub[0] and subbeg[nbins] (number of
// subproblems).  (A last-block scan inside k_colscan would need a device-scope
// fence in every block, which writes back L2 and costs more than this launch.)
constexpr int kScanThreads = 1024;
__device__ __forceinline__ void scan_groups_body(int2* __restrict__ gsum, int ngroups, int nbins,
                                                 int* __restrict__ subbeg, int* __restrict__ nsub)
{
    __shared__ int w_pts[kScanThreads / 64], w_sub[kScanThreads / 64];
    __shared__ int carry[2];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int cpts = 0, csub = 0;
    for (int base = 0; base < ngroups; base += kScanThreads) {
        const int g = base + t;
        const int2 v = g < ngroups ? gsum[g] : make_int2(0, 0);
        int ia = v.x, ib = v.y;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int na = __shfl_up(ia, o, 64), nb = __shfl_up(ib, o, 64);
            if (lane >= o) { ia += na; ib += nb; }
        }
        if (lane == 63) { w_pts[w] = ia; w_sub[w] = ib; }
        __syncthreads();
        int pa = cpts + ia - v.x, pb = csub + ib - v.y;
        for (int k = 0; k < w; ++k) { pa += w_pts[k]; pb += w_sub[k]; }
        if (g < ngroups) gsum[g] = make_int2(pa, pb);
        if (t == kScanThreads - 1) { carry[0] = pa + v.x; carry[1] = pb + v.y; }
        __syncthreads();
        cpts = carry[0];
        csub = carry[1];
        __syncthreads();
    }
    if (t == 0) {
        *nsub = csub;
        subbeg[nbins] = csub;
    }
}
__global__ __launch_bounds__(kScanThreads) void k_scan_groups(int2* __restrict__ gsum, int ngroups, int nbins,
                                                              int* __restrict__ subbeg, int* __restrict__ nsub)
{
    scan_groups_body(gsum, ngroups, nbins, subbeg, nsub);
}

// Pass 3: scatter the points to their sorted positions, storing what the
// gather and spread need: the fractional offsets t = c x - rint(c x) (exact),
// the packed cell (mu2 << 16 | mu1) and the original index.  The first nbins
// threads also write their bin's first subproblem (subbeg) and subproblems
// (bin, start, count <= maxpts), in bin order.  A bin starts at
// gsum[bin / kColBins].x + gexcl[bin].x.
template <typename T>
__device__ __forceinline__ void scatter_body(int blk, const T* __restrict__ x, const T* __restrict__ y, int M, int chunk,
                                             int L1, int L2, T inv1, T inv2, int nb1, int nbins, int maxpts,
                                             const int* __restrict__ off, const int* __restrict__ tot,
                                             const int2* __restrict__ gexcl, const int2* __restrict__ gsum,
                                             const int* __restrict__ rank, PtRec<T>* __restrict__ pts,
                                             int* __restrict__ subbeg, int4* __restrict__ sub)
{
    const int i = blk * 256 + threadIdx.x;
    if (i < nbins) {
        const int c = tot[i];
        const int2 e = gexcl[i], gb = gsum[i / kColBins];
        const int pos = gb.x + e.x, sp = gb.y + e.y;
        subbeg[i] = sp;
        const int ns = (c + maxpts - 1) / maxpts;
        for (int k = 0; k < ns; ++k) sub[sp + k] = make_int4(i, pos + k * maxpts, min(maxpts, c - k * maxpts), 0);
    }
    if (i >= M) return;
    int m1, m2;
    T t1, t2;
    const int bin = bin_of(x[i], y[i], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
    const int p = gsum[bin / kColBins].x + gexcl[bin].x + off[(size_t)(i / chunk) * nbins + bin] + rank[i];
    PtRec<T> rec;
    rec.t1 = t1;
    rec.t2 = t2;
    rec.cell = (m2 << 16) | m1;
    rec.idx = i;
    pts[p] = rec;
}
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const T* __restrict__ x, const T* __restrict__ y, int M, int chunk,
                                                 int L1, int L2, T inv1, T inv2, int nb1, int nbins, int maxpts,
                                                 const int* __restrict__ off, const int* __restrict__ tot,
                                                 const int2* __restrict__ gexcl, const int2* __restrict__ gsum,
                                                 const int* __restrict__ rank, PtRec<T>* __restrict__ pts,
                                                 int* __restrict__ subbeg, int4* __restrict__ sub)
{
    scatter_body<T>(blockIdx.x, x, y, M, chunk, L1, L2, inv1, inv2, nb1, nbins, maxpts, off, tot, gexcl, gsum, rank,
                    pts, subbeg, sub);
}

// Pass 3 with the group scan folded in (at most kScatThreads groups of
// kColBins bins, i.e. 32,768 bins): every block scans the group sums gsum
// itself into LDS (at most 8 KiB of loads per block), so no launch separates
// k_colscan from the scatter.  A block scatters kScatThreads * kScatPer points
// and, if its bin range [blk * kScatThreads, ...) holds bins, writes their
// subproblems.  Same positions as k_scan_groups + k_scatter.
constexpr int kScatThreads = 1024, kScatPer = 4;
template <typename T>
__global__ __launch_bounds__(kScatThreads) void k_scatter_scan(
    const T* __restrict__ x, const T* __restrict__ y, int M, int chunk, int L1, int L2, T inv1, T inv2, int nb1,
    int nbins, int maxpts, const int* __restrict__ off, const int* __restrict__ tot, const int2* __restrict__ gexcl,
    const int2* __restrict__ gsum, int ngroups, const int* __restrict__ rank, PtRec<T>* __restrict__ pts,
    int* __restrict__ subbeg, int4* __restrict__ sub, int* __restrict__ nsub)
{
    __shared__ int2 pre[kScatThreads + 1];
    __shared__ int w_a[kScatThreads / 64], w_b[kScatThreads / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // this block's points: loads in flight during the group scan
    const int p0 = blockIdx.x * (kScatThreads * kScatPer);
    T xv[kScatPer], yv[kScatPer];
    int rv[kScatPer];
#pragma unroll
    for (int q = 0; q < kScatPer; ++q) {
        const int i = p0 + q * kScatThreads + t;
        const bool in = i < M;
        xv[q] = in ? x[i] : (T)0;
        yv[q] = in ? y[i] : (T)0;
        rv[q] = in ? rank[i] : 0;
    }
    // exclusive scan of the (points, subproblems) group sums, one per thread
    const int2 v = t < ngroups ? gsum[t] : make_int2(0, 0);
    int ia = v.x, ib = v.y;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int na = __shfl_up(ia, o, 64), nb = __shfl_up(ib, o, 64);
        if (lane >= o) { ia += na; ib += nb; }
    }
    if (lane == 63) { w_a[w] = ia; w_b[w] = ib; }
    __syncthreads();
    int pa = ia - v.x, pb = ib - v.y;
    for (int k = 0; k < w; ++k) { pa += w_a[k]; pb += w_b[k]; }
    pre[t] = make_int2(pa, pb);
    if (t == kScatThreads - 1) pre[kScatThreads] = make_int2(pa + v.x, pb + v.y);
    __syncthreads();
    const int bi = blockIdx.x * kScatThreads + t;
    if (bi < nbins) {
        const int c = tot[bi];
        const int2 e = gexcl[bi], gb = pre[bi / kColBins];
        const int pos = gb.x + e.x, sp = gb.y + e.y;
        subbeg[bi] = sp;
        const int ns = (c + maxpts - 1) / maxpts;
        for (int k = 0; k < ns; ++k) sub[sp + k] = make_int4(bi, pos + k * maxpts, min(maxpts, c - k * maxpts), 0);
    }
    if (blockIdx.x == 0 && t == 0) {
        const int total = pre[kScatThreads].y;
        *nsub = total;
        subbeg[nbins] = total;
    }
#pragma unroll
    for (int q = 0; q < kScatPer; ++q) {
        const int i = p0 + q * kScatThreads + t;
        if (i < M) {
            int m1, m2;
            T t1, t2;
            const int bin = bin_of(xv[q], yv[q], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
            const int p = pre[bin / kColBins].x + gexcl[bin].x + off[(size_t)(i / chunk) * nbins + bin] + rv[q];
            PtRec<T> rec;
            rec.t1 = t1;
            rec.t2 = t2;
            rec.cell = (m2 << 16) | m1;
            rec.idx = i;
            pts[p] = rec;
        }
    }
}

// Pass 3 for grids whose bins fit one LDS array (nbins <= kMaxLdsBins): block
// b takes the points of histogram block b, [b*chunk, (b+1)*chunk), and builds
// base[bin] = first sorted position of (block b, bin) in LDS from its row of
// column offsets.  Each point then takes base[bin] + its rank by incrementing
// base[bin] (one LDS atomic per run of equal bins in a wave), so pass 1 stores
// no per-point rank.  Inside a (block, bin) the points are ranked in the order
// the waves reach the counter, as in pass 1.  The bins' subproblem lists are
// split evenly over the blocks.
static_assert(kMaxLdsBins / kColBins <= kScatThreads, "k_scatter_lds scans at most kScatThreads groups");
template <typename T>
__global__ __launch_bounds__(kScatThreads) void k_scatter_lds(
    const T* __restrict__ x, const T* __restrict__ y, int M, int chunk, int L1, int L2, T inv1, T inv2, int nb1,
    int nbins, int maxpts, const int* __restrict__ off, const int* __restrict__ tot, const int2* __restrict__ gexcl,
    const int2* __restrict__ gsum, int ngroups, PtRec<T>* __restrict__ pts, int* __restrict__ subbeg,
    int4* __restrict__ sub, int* __restrict__ nsub)
{
    __shared__ int base[kMaxLdsBins];
    __shared__ int2 pre[kScatThreads + 1];
    __shared__ int w_a[kScatThreads / 64], w_b[kScatThreads / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lo = blockIdx.x * chunk, hi = min(lo + chunk, M);
    // the first batch of points: loads in flight during the scan
    T xv[kScatPer], yv[kScatPer];
#pragma unroll
    for (int q = 0; q < kScatPer; ++q) {
        const int i = lo + q * kScatThreads + t;
        xv[q] = i < hi ? x[i] : (T)0;
        yv[q] = i < hi ? y[i] : (T)0;
    }
    const int* __restrict__ orow = off + (size_t)blockIdx.x * nbins;
    // exclusive scan of the (points, subproblems) group sums, one per thread
    const int2 v = t < ngroups ? gsum[t] : make_int2(0, 0);
    int ia = v.x, ib = v.y;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int na = __shfl_up(ia, o, 64), nb = __shfl_up(ib, o, 64);
        if (lane >= o) { ia += na; ib += nb; }
    }
    if (lane == 63) { w_a[w] = ia; w_b[w] = ib; }
    __syncthreads();
    int pa = ia - v.x, pb = ib - v.y;
    for (int k = 0; k < w; ++k) { pa += w_a[k]; pb += w_b[k]; }
    pre[t] = make_int2(pa, pb);
    if (t == kScatThreads - 1) pre[kScatThreads] = make_int2(pa + v.x, pb + v.y);
    __syncthreads();
    for (int b = t; b < nbins; b += kScatThreads) base[b] = pre[b / kColBins].x + gexcl[b].x + orow[b];
    // subproblem lists: an even share of the bins per block
    const int bpb = (nbins + (int)gridDim.x - 1) / (int)gridDim.x;
    const int blo = blockIdx.x * bpb, bhi = min(blo + bpb, nbins);
    for (int bi = blo + t; bi < bhi; bi += kScatThreads) {
        const int c = tot[bi];
        const int2 e = gexcl[bi], gb = pre[bi / kColBins];
        const int pos = gb.x + e.x, sp = gb.y + e.y;
        subbeg[bi] = sp;
        const int ns = (c + maxpts - 1) / maxpts;
        for (int k = 0; k < ns; ++k) sub[sp + k] = make_int4(bi, pos + k * maxpts, min(maxpts, c - k * maxpts), 0);
    }
    if (blockIdx.x == 0 && t == 0) {
        const int total = pre[kScatThreads].y;
        *nsub = total;
        subbeg[nbins] = total;
    }
    __syncthreads();
    for (int b0 = lo; b0 < hi; b0 += kScatPer * kScatThreads) {
        if (b0 != lo) {
#pragma unroll
            for (int q = 0; q < kScatPer; ++q) {
                const int i = b0 + q * kScatThreads + t;
                xv[q] = i < hi ? x[i] : (T)0;
                yv[q] = i < hi ? y[i] : (T)0;
            }
        }
#pragma unroll
        for (int q = 0; q < kScatPer; ++q) {
            const int i = b0 + q * kScatThreads + t;
            const bool act = i < hi;
            int m1, m2;
            T t1, t2;
            const int bin = bin_of(xv[q], yv[q], L1, L2, inv1, inv2, nb1, m1, m2, t1, t2);
            const int p = wave_run_rank(base, bin, act);
            if (act) {
                PtRec<T> rec;
                rec.t1 = t1;
                rec.t2 = t2;
                rec.cell = (m2 << 16) | m1;
                rec.idx = i;
                pts[p] = rec;
            }
        }
    }
}

// ================================================================= window taps
struct WinArgs {
    WinPoly val;    // phi_hat / S
    WinPoly der;    // phi_hat' / S
    const WinTables64* __restrict__ w64;   // FP64 plans: degree-14 tables in device memory
};

__device__ __forceinline__ void taps(float t, const WinArgs& wa, float (&w)[kW]) { window_taps(t, wa.val, w); }
__device__ __forceinline__ void taps(double t, const WinArgs& wa, double (&w)[kW])
{
    window_taps64(t, wa.w64->val, w);
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
    window_dtaps64(t, wa.w64->der, w);
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

// LDS row stride of a gather tile (complex elements).  40: consecutive rows
// start 16 banks apart (with 32 every row starts at the same bank); polar
// 512^2 gather 31.6 -> 30.7 us (36: 30.5, 44: 30.3).
#ifndef NFFT2D_TILE_LD
#define NFFT2D_TILE_LD 40
#endif
constexpr int kTileLd = NFFT2D_TILE_LD;

// Launched with 256 threads; the 784 loads of a thread group are issued
// together (a thread has at most 4): tile_fetch into registers, tile_store
// from them into LDS (split so that a persistent kernel can fetch the next
// tile while it works on the current one).
constexpr int kTileW = kBin + 2 * kK;                       // 28 used rows / columns
constexpr int kTilePer = (kTileW * kTileW + 255) / 256;     // tile values per thread
template <typename C>
__device__ __forceinline__ void tile_fetch(C (&v)[kTilePer], const C* __restrict__ Gt, int L1, int L2, int o1, int o2)
{
    constexpr int W = kTileW;
    const bool big = L1 >= W && L2 >= W;
#pragma unroll
    for (int q = 0; q < kTilePer; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        if (e < W * W) {
            const int r = e / W, c = e - r * W;
            const int g2 = big ? wrap1(o2 + r, L2) : wrap(o2 + r, L2);
            const int g1 = big ? wrap1(o1 + c, L1) : wrap(o1 + c, L1);
            v[q] = Gt[(size_t)g2 * L1 + g1];
        }
    }
}
template <typename C>
__device__ __forceinline__ void tile_store(C* tile, const C (&v)[kTilePer])
{
    constexpr int W = kTileW;
#pragma unroll
    for (int q = 0; q < kTilePer; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        if (e < W * W) {
            const int r = e / W, c = e - r * W;
            tile[r * kTileLd + c] = v[q];
        }
    }
}
template <typename C>
__device__ __forceinline__ void load_tile(C* tile, const C* __restrict__ Gt, int L1, int L2, int o1, int o2)
{
    C v[kTilePer];
    tile_fetch(v, Gt, L1, L2, o1, o2);
    tile_store(tile, v);
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

// (c) type-2 gather over one subproblem (<= 256 points of one bin), one point
// per thread: the complex128 gathers and the derivative modes of both
// precisions (the complex64 value gather is k_gather_pairs, below).
//   MODE 0: out[l] = y_l;
//   MODE 1: out = y, out_d1 = dy/dx1, out_d2 = dy/dx2;
//   MODE 2: gradient Re(conj(g_l) dy_l/dx) into gx1 / gx2;
//   MODE 3: least squares, g = y_obs: out = r = 2 (y - y_obs), gx = Re(conj(r) dy/dx);
//           out_d1 receives r once more in the plan's sorted point order
//           (out_d1[p] for the point record p), for a spread that reads its
//           samples in that order; out may be null, out_d1 may not.
// y_l = sum_{m2} w2[m2] sum_{m1} w1[m1] G[(mu2+m2-K) mod L2][(mu1+m1-K) mod L1]
// (BP_FBP_FFBP/NFFT1_2D.m:61-67 on the unshifted grid).  The block first
// counting-sorts its points by cell in LDS so that neighbouring lanes read
// neighbouring tile addresses.
template <typename T, int MODE>
__device__ __forceinline__ void gather_body(int s, const cplx<T>* __restrict__ Gt, const BinGeom& geo,
                                            const PtRec<T>* __restrict__ pts, const int4* __restrict__ sub,
                                            const WinArgs& wa, cplx<T>* __restrict__ out, cplx<T>* __restrict__ out_d1,
                                            cplx<T>* __restrict__ out_d2, const cplx<T>* __restrict__ g,
                                            T* __restrict__ gx1, T* __restrict__ gx2)
{
    using C = cplx<T>;
    constexpr int W = kTileLd;
    __shared__ C tile[(kBin + 2 * kK) * kTileLd];
    __shared__ int cnt[kBin * kBin];
    __shared__ int s_warp[kGatherThreads / 64];
    __shared__ T s_f1[kGatherThreads], s_f2[kGatherThreads];
    __shared__ int s_cell[kGatherThreads], s_idx[kGatherThreads];
    __shared__ int s_pos[MODE == 3 ? kGatherThreads : 1];
    const int4 d = sub[s];
    const int b1 = d.x % geo.nb1, b2 = d.x / geo.nb1;
    const int O1 = b1 * kBin, O2 = b2 * kBin;
    const int tid = threadIdx.x;
    const bool have = tid < d.z;
    int cell = 0, o = 0;
    T f1 = 0, f2 = 0;
    if (have) {
        const PtRec<T> q = pts[d.y + tid];
        const int pc = q.cell;
        o = q.idx;
        f1 = q.t1;
        f2 = q.t2;
        cell = min(max((pc >> 16) - O2, 0), kBin - 1) * kBin + min(max((pc & 0xffff) - O1, 0), kBin - 1);
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
        if constexpr (MODE == 3) s_pos[dst] = tid;
    }
    __syncthreads();
    if (!have) return;         // (no barrier follows in this function)
    f1 = s_f1[tid];
    f2 = s_f2[tid];
    cell = s_cell[tid];
    o = s_idx[tid];
    const int ps = MODE == 3 ? d.y + s_pos[tid] : 0;   // this point's record position
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
        } else if constexpr (MODE == 2) {
            const C gg = g[o];
            gx1[o] = cc * (gg.x * bx + gg.y * by);
            gx2[o] = cc * (gg.x * cx + gg.y * cy);
        } else {                       // MODE 3: g holds y_obs; r = 2 (y - y_obs) and its MODE 2 gradient
            const C yo = g[o];
            const C r = mk((T)2 * (ax - yo.x), (T)2 * (ay - yo.y));
            // both stores unconditional (a branch around them makes the
            // compiler hold the whole tile window in registers); without a
            // caller's r the first one repeats the second
            *(out ? out + o : out_d1 + ps) = r;
            out_d1[ps] = r;
            gx1[o] = cc * (r.x * bx + r.y * by);
            gx2[o] = cc * (r.x * cx + r.y * cy);
        }
    }
}

// (c) type-2 gather, complex64, MODE 0, with up to TWO points per lane.
// The tile reads are what binds the gather (a lane reads 13 x 13 tile values
// per point from LDS; the FMAs have room to spare), so a lane takes two points
// of one cell pair (cells 2j and 2j+1 of a cell row of the bin) and reads
// their union window once: 13 rows of 14 values serve both points, 182 reads
// instead of 338, as 16-byte reads (the window starts at an even column).
// The block counting-sorts its points by cell pair; a pair with n points takes
// ceil(n / 2) lanes.  Each point's 13 column weights are placed in a row of 14
// with a zero at the end its window does not reach, so a point's sum runs in
// the same order as in gather_body (plus one FMA with a zero weight):
// the results are bitwise the same.
constexpr int kPairs = kBin * kBin / 2;                     // cell pairs of a bin
// A block takes one subproblem (NSUB = 1) or, with NSUB = 2, also the last
// subproblem of the same bin when that one holds at most kTailMax points.  A
// cell pair with n points takes ceil(n / 2) lanes, so at most (points +
// kPairs) / 2 lanes in all: one lane per thread for up to 384 points.
constexpr int kTailMax = kGatherThreads / 2;
template <int NSUB> constexpr int pair_lanes() { return (kGatherThreads + (NSUB - 1) * kTailMax + kPairs) / 2; }
static_assert(pair_lanes<2>() <= kGatherThreads, "at most one lane slot per thread");
template <int NSUB>
struct __attribute__((aligned(16))) PairLds {
    float2 tile[kTileW * kTileLd];            // 16-byte aligned: the window reads are 16-byte loads
    int cnt[kPairs];
    int s_warp[kGatherThreads / 64];
    float t1[2][pair_lanes<NSUB>()], t2[2][pair_lanes<NSUB>()];
    int cell[2][pair_lanes<NSUB>()], idx[2][pair_lanes<NSUB>()];
    int nlanes;
};

// Sort the block's points (thread t holds point t of subproblem q in rec[q]
// when t < n[q]) by cell pair into the lane slots of L; returns the number of
// lanes.  The tile may be stored into L before the call (the first barrier
// publishes it).
template <int NSUB>
__device__ __forceinline__ int pairs_stage(PairLds<NSUB>& L, const int (&n)[NSUB], int O1, int O2,
                                           const PtRec<float> (&rec)[NSUB])
{
    const int tid = threadIdx.x;
    int cell[NSUB], rk[NSUB];
#pragma unroll
    for (int q = 0; q < NSUB; ++q)
        cell[q] = tid < n[q] ? min(max((rec[q].cell >> 16) - O2, 0), kBin - 1) * kBin +
                                   min(max((rec[q].cell & 0xffff) - O1, 0), kBin - 1)
                             : 0;                            // cell pair: cell >> 1 = row * 8 + column / 2
    if (tid < kPairs) L.cnt[tid] = 0;
    for (int l = tid; l < pair_lanes<NSUB>(); l += kGatherThreads) L.idx[1][l] = -1;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NSUB; ++q) rk[q] = tid < n[q] ? atomicAdd(&L.cnt[cell[q] >> 1], 1) : 0;
    __syncthreads();
    const int lanes = tid < kPairs ? (L.cnt[tid] + 1) >> 1 : 0;
    const int start = block_exclusive_scan(lanes, L.s_warp);
    __syncthreads();
    if (tid < kPairs) L.cnt[tid] = start;
    if (tid == kGatherThreads - 1) L.nlanes = start + lanes;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NSUB; ++q) {
        if (tid < n[q]) {
            const int l = L.cnt[cell[q] >> 1] + (rk[q] >> 1), k = rk[q] & 1;
            L.t1[k][l] = rec[q].t1;
            L.t2[k][l] = rec[q].t2;
            L.cell[k][l] = cell[q];
            L.idx[k][l] = rec[q].idx;
        }
    }
    __syncthreads();
    return L.nlanes;
}

// The 13 x 13 sums of lane l's one or two points (L: any layout with the
// tile and the slot arrays of PairLds).
template <class Lds>
__device__ __forceinline__ void pairs_mac(const Lds& L, int l, const WinArgs& wa, float2* __restrict__ out)
{
    constexpr int W = kTileLd;
    const int ca = L.cell[0][l], oa = L.idx[0][l];
    const int ob = L.idx[1][l];
    const bool two = ob >= 0;
    const int cb = two ? L.cell[1][l] : ca;
    const float fa1 = L.t1[0][l], fa2 = L.t2[0][l];
    const float fb1 = two ? L.t1[1][l] : fa1, fb2 = two ? L.t2[1][l] : fa2;
    const int r0 = ca / kBin, c0 = (ca % kBin) & ~1;          // window: rows r0.., columns c0..c0+13
    const bool sa = (ca & 1) != 0, sb = (cb & 1) != 0;       // point in the odd cell: window shifted by one
    float wa1[kW], wa2[kW], wb1[kW], wb2[kW];
    taps(fa1, wa, wa1);
    taps(fa2, wa, wa2);
    taps(fb1, wa, wb1);
    taps(fb2, wa, wb2);
    float ea[kW + 1], eb[kW + 1];
#pragma unroll
    for (int j = 0; j <= kW; ++j) {
        const float a0 = j < kW ? wa1[j] : 0.f, a1 = j > 0 ? wa1[j - 1] : 0.f;
        const float b0 = j < kW ? wb1[j] : 0.f, b1 = j > 0 ? wb1[j - 1] : 0.f;
        ea[j] = sa ? a1 : a0;
        eb[j] = sb ? b1 : b0;
    }
    const float2* base = L.tile + r0 * W + c0;
    float xa = 0.f, ya = 0.f, xb = 0.f, yb = 0.f;
#pragma unroll
    for (int b = 0; b < kW; ++b) {
        float sxa = 0.f, sya = 0.f, sxb = 0.f, syb = 0.f;
#pragma unroll
        for (int j = 0; j <= kW; ++j) {
            const float2 u = base[b * W + j];
            sxa = fmaf(ea[j], u.x, sxa);
            sya = fmaf(ea[j], u.y, sya);
            sxb = fmaf(eb[j], u.x, sxb);
            syb = fmaf(eb[j], u.y, syb);
        }
        xa = fmaf(wa2[b], sxa, xa);
        ya = fmaf(wa2[b], sya, ya);
        xb = fmaf(wb2[b], sxb, xb);
        yb = fmaf(wb2[b], syb, yb);
    }
    // A lane with one point carries it in both slots (same inputs, so the same
    // bits): the second store is unconditional, which keeps the compiler from
    // sinking the second point's sums into a branch (it would then hold every
    // tile row until that branch).
    out[oa] = make_float2(xa, ya);
    out[two ? ob : oa] = make_float2(xb, yb);
}

// One point of a subproblem of the bin with origin (O1, O2), its 13 x 14
// window read from the grid instead of an LDS tile: the fmaf sequence of
// pairs_mac for a point (columns c0..c0+13 with the weights shifted for an odd
// cell, then the 13 rows), so the same bits.  For subproblems of a few points,
// where a tile, a sort and a wave's sums per subproblem cost more than the
// 182 cached reads per point.
__device__ __forceinline__ float2 direct_point(const PtRec<float>& rec, int O1, int O2, const float2* __restrict__ Gt,
                                               const BinGeom& geo, const WinArgs& wa)
{
    const int cell = min(max((rec.cell >> 16) - O2, 0), kBin - 1) * kBin + min(max((rec.cell & 0xffff) - O1, 0), kBin - 1);
    const int r0 = cell / kBin, c0 = (cell % kBin) & ~1;
    const bool odd = (cell & 1) != 0;
    float w1[kW], w2[kW];
    taps(rec.t1, wa, w1);
    taps(rec.t2, wa, w2);
    float e[kW + 1];
#pragma unroll
    for (int j = 0; j <= kW; ++j) {
        const float a0 = j < kW ? w1[j] : 0.f, a1 = j > 0 ? w1[j - 1] : 0.f;
        e[j] = odd ? a1 : a0;
    }
    const bool big = geo.L1 >= kTileW && geo.L2 >= kTileW;
    int g1[kW + 1];
#pragma unroll
    for (int j = 0; j <= kW; ++j) {
        const int o1 = O1 - kK + c0 + j;
        g1[j] = big ? wrap1(o1, geo.L1) : wrap(o1, geo.L1);
    }
    // one row (14 loads) in flight at a time; the row weights shift down so
    // that the current one is w2[0] (register arrays take static indices only)
    float x = 0.f, y = 0.f;
#pragma unroll 1
    for (int b = 0; b < kW; ++b) {
        const int o2 = O2 - kK + r0 + b;
        const float2* __restrict__ row = Gt + (size_t)(big ? wrap1(o2, geo.L2) : wrap(o2, geo.L2)) * geo.L1;
        float sx = 0.f, sy = 0.f;
#pragma unroll
        for (int j = 0; j <= kW; ++j) {
            const float2 u = row[g1[j]];
            sx = fmaf(e[j], u.x, sx);
            sy = fmaf(e[j], u.y, sy);
        }
        x = fmaf(w2[0], sx, x);
        y = fmaf(w2[0], sy, y);
#pragma unroll
        for (int m = 0; m + 1 < kW; ++m) w2[m] = w2[m + 1];
    }
    return make_float2(x, y);
}

// NSUB = 2: a bin's last subproblem with at most kTailMax points is taken by
// the block of the subproblem before it (which holds more), and its own block
// returns at once.  The two share one tile and one staging pass instead of
// paying both for a small tail (polar 512^2 / M = 1M: 4,545 blocks with points
// instead of 5,456).
template <int NSUB>
__device__ __forceinline__ void gather_pairs_body(int s, int ns, const float2* __restrict__ Gt, const BinGeom& geo,
                                                  const PtRec<float>* __restrict__ pts, const int4* __restrict__ sub,
                                                  const WinArgs& wa, float2* __restrict__ out)
{
    __shared__ PairLds<NSUB> L;
    const int4 d = sub[s];
    int n[NSUB], at[NSUB];
    n[0] = d.z;
    at[0] = d.y;
    if constexpr (NSUB == 2) {
        // entries of a bin are consecutive, all but its last hold maxpts > kTailMax points
        const int4 a = s > 0 ? sub[s - 1] : make_int4(-1, 0, 0, 0);
        const int4 e = s + 1 < ns ? sub[s + 1] : make_int4(-1, 0, 0, 0);
        if (d.z <= kTailMax && a.x == d.x && a.z > kTailMax) return;
        n[1] = e.x == d.x && e.z <= kTailMax && d.z > kTailMax ? e.z : 0;
        at[1] = e.y;
    }
    const int O1 = (d.x % geo.nb1) * kBin, O2 = (d.x / geo.nb1) * kBin;
    PtRec<float> rec[NSUB];
#pragma unroll
    for (int q = 0; q < NSUB; ++q) {
        rec[q].t1 = rec[q].t2 = 0.f;
        rec[q].cell = rec[q].idx = 0;
        if ((int)threadIdx.x < n[q]) rec[q] = pts[at[q] + threadIdx.x];
    }
    load_tile(L.tile, Gt, geo.L1, geo.L2, O1 - kK, O2 - kK);
    const int nl = pairs_stage<NSUB>(L, n, O1, O2, rec);
    if ((int)threadIdx.x < nl) pairs_mac(L, threadIdx.x, wa, out);
}

#ifndef NFFT2D_GATHER_WAVES
#define NFFT2D_GATHER_WAVES 6
#endif
// The complex64 type-2 gather of a plan: one block per subproblem, with
// small tails merged (NSUB = 2) or not (NSUB = 1).
template <int NSUB>
__global__ __launch_bounds__(kGatherThreads) __attribute__((amdgpu_waves_per_eu(NFFT2D_GATHER_WAVES, 8))) void
k_gather_pairs(const float2* __restrict__ Gt, BinGeom geo, const PtRec<float>* __restrict__ pts,
               const int4* __restrict__ sub, const int* __restrict__ nsub, WinArgs wa, float2* __restrict__ out)
{
    const int ns = *nsub;
    const int s = xcd_runs(blockIdx.x);   // runs of neighbouring subproblems share an L2
    if (s >= ns) return;
    gather_pairs_body<NSUB>(s, ns, Gt, geo, pts, sub, wa, out);
}

template <typename T, int MODE>
__global__ __launch_bounds__(kGatherThreads) void k_gather(
    const cplx<T>* __restrict__ Gt, BinGeom geo, const PtRec<T>* __restrict__ pts,
    const int4* __restrict__ sub, const int* __restrict__ nsub, WinArgs wa, cplx<T>* __restrict__ out,
    cplx<T>* __restrict__ out_d1, cplx<T>* __restrict__ out_d2, const cplx<T>* __restrict__ g, T* __restrict__ gx1,
    T* __restrict__ gx2)
{
    if ((int)blockIdx.x >= *nsub) return;
    gather_body<T, MODE>(blockIdx.x, Gt, geo, pts, sub, wa, out, out_d1, out_d2, g, gx1, gx2);
}

// LDS writes of a wave become visible to the whole wave.
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (c) type-2 gather, one WAVE per subproblem (complex64, MODE 0), for the
// persistent batch gather.  Many subproblems are in flight per CU (four per
// block), which pays off when they are small (sparse point sets).  Same
// arithmetic per point as k_gather_pairs, so the same bits.

// One point's 13 x 13 sum from a bin's LDS tile (row stride W): the fmaf
// sequence of pairs_mac for a lane's first point (its 13 column weights in a
// row of 14 with a zero at the end its window does not reach, then the 13
// rows), so the same bits.
template <int W>
__device__ __forceinline__ float2 tile_point(const float2* tile, const PtRec<float>& rec, int O1, int O2,
                                             const WinArgs& wa)
{
    const int cell = min(max((rec.cell >> 16) - O2, 0), kBin - 1) * kBin + min(max((rec.cell & 0xffff) - O1, 0), kBin - 1);
    const int r0 = cell / kBin, c0 = (cell % kBin) & ~1;
    const bool odd = (cell & 1) != 0;
    float w1[kW], w2[kW];
    taps(rec.t1, wa, w1);
    taps(rec.t2, wa, w2);
    float e[kW + 1];
#pragma unroll
    for (int j = 0; j <= kW; ++j) {
        const float a0 = j < kW ? w1[j] : 0.f, a1 = j > 0 ? w1[j - 1] : 0.f;
        e[j] = odd ? a1 : a0;
    }
    const float2* base = tile + r0 * W + c0;
    float x = 0.f, y = 0.f;
#pragma unroll
    for (int b = 0; b < kW; ++b) {
        float sx = 0.f, sy = 0.f;
#pragma unroll
        for (int j = 0; j <= kW; ++j) {
            const float2 u = base[b * W + j];
            sx = fmaf(e[j], u.x, sx);
            sy = fmaf(e[j], u.y, sy);
        }
        x = fmaf(w2[b], sx, x);
        y = fmaf(w2[b], sy, y);
    }
    return make_float2(x, y);
}

// The persistent batch gather's subproblem, one point per lane: the wave
// stages the 28 x 28 tile in its own LDS tile (13 loads per lane, all in
// flight), then each lane sums its own point from it, 64 points per pass.  A
// pass of the pairs_mac form (a wave-level counting sort by cell pair, two
// points per lane) also covered at most 64 points and issued the same window
// reads per lane, so pairing bought no pass and no read here; on sparse sets
// (mostly one point per cell pair) it ran two points' FMAs per lane for one
// point: a batch of 256 uneven transforms (16M points) took 1.78 ms per step
// with it, 1.60 without.
// Its tile rows are kWaveTileLd complex values apart: with 32 every row
// starts at the same LDS bank, so the windows of a 16-lane read group (one
// point per lane, at scattered cells) fell on at most 8 distinct 16-byte bank
// groups.  36: each row starts 8 banks (of 64) further; that batch 1.60 -> 1.44 ms per step
// (40 and 44 the same, 48 1.75 ms: the larger tiles cost a block per CU).
#ifndef NFFT2D_WAVE_TILE_LD
#define NFFT2D_WAVE_TILE_LD 36
#endif
constexpr int kWaveTileLd = NFFT2D_WAVE_TILE_LD;
struct __attribute__((aligned(16))) WaveTileLds {
    float2 tile[kTileW * kWaveTileLd];
};
__device__ __forceinline__ void gather_wave_single(const int4& d, WaveTileLds& L, const float2* __restrict__ Gt,
                                                   const BinGeom& geo, const PtRec<float>* __restrict__ pts,
                                                   const WinArgs& wa, float2* __restrict__ out)
{
    constexpr int W = kTileW;
    constexpr int kPer = (W * W + 63) / 64;                  // 13
    const int lane = threadIdx.x & 63;
    const int O1 = (d.x % geo.nb1) * kBin, O2 = (d.x / geo.nb1) * kBin;
    auto fetch = [&](int k) {
        PtRec<float> r;
        r.t1 = r.t2 = 0.f;
        r.cell = r.idx = 0;
        if (k < d.z) r = pts[d.y + k];
        return r;
    };
    // the first pass's records and the tile, all loads in flight
    PtRec<float> rec = fetch(lane);
    float2 v[kPer];
    const bool big = geo.L1 >= W && geo.L2 >= W;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = lane + 64 * q;
        v[q] = make_float2(0.f, 0.f);
        if (e < W * W) {
            const int r = e / W, c = e - r * W;
            const int o2 = O2 - kK + r, o1 = O1 - kK + c;
            const int g2 = big ? wrap1(o2, geo.L2) : wrap(o2, geo.L2);
            const int g1 = big ? wrap1(o1, geo.L1) : wrap(o1, geo.L1);
            v[q] = Gt[(size_t)g2 * geo.L1 + g1];
        }
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = lane + 64 * q;
        if (e < W * W) {
            const int r = e / W, c = e - r * W;
            L.tile[r * kWaveTileLd + c] = v[q];
        }
    }
    wave_sync_lds();
    for (int base = 0; base < d.z; base += 64) {
        const PtRec<float> nxt = fetch(base + 64 + lane);    // the next pass's, in flight during this one
        if (base + lane < d.z) out[rec.idx] = tile_point<kWaveTileLd>(L.tile, rec, O1, O2, wa);
        rec = nxt;
    }
    wave_sync_lds();                                         // the tile is reused by the next subproblem
}

// (d) type-1 spread (interim form): each block owns one subproblem, accumulates
// the bin + K-halo in LDS, then adds it into the grid with global atomics.
// SORTED: data[p] is the sample of point record p (the plan's sorted order)
// instead of data[original index].
template <typename T, bool SORTED>
__global__ __launch_bounds__(256) void k_spread(cplx<T>* __restrict__ Gt, BinGeom geo, const PtRec<T>* __restrict__ pts,
                                                const cplx<T>* __restrict__ data,
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
        const PtRec<T> q = pts[p];
        const int pc = q.cell;
        const C v = data[SORTED ? p : q.idx];
        const T f1 = q.t1, f2 = q.t2;
        const int c0 = min(max((pc & 0xffff) - O1, 0), kBin - 1);
        const int r0 = min(max((pc >> 16) - O2, 0), kBin - 1);
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
// (d) type-1 spread, point-stationary.  Each subproblem (<= 256 points of one
// 16x16 bin) accumulates its points' footprints into a 32x32 tile whose origin
// is the bin origin minus the K = 6 halo (rows / columns 0..27 are used):
//     T[i][j] = sum_p (d_p w2_p[i - r_p]) * w1_p[j - c_p],   r_p, c_p = cell - bin origin.
// Every point is an outer product of two 32-vectors, and
// v_mfma_f32_32x32x16_f16 adds 16 of them per instruction into a 32x32 FP32
// accumulator held in registers.  Each operand value x is split x = hi + lo
// into two FP16 values (round to nearest) and T accumulates
// lo*hi + hi*lo + hi*hi, which keeps about 22 bits of every product.  A wave
// scales its data values by a power of two that puts their largest component
// in [2^14, 2^15) (the window taps are below 1, so every operand stays inside
// the FP16 range) and scales its tile back before the 4 waves' tiles are
// summed in a fixed order.  The tiles go to scratch, and k_spread_merge sums,
// for every grid cell, the tiles of the (up to 9) neighbouring bins that
// cover it, in a fixed order.  No atomics; needs at least 3 bins and L >= 44
// per dimension (the plan falls back to the LDS-atomic k_spread otherwise).
// Operand staging, 16 points at a time: a point's 13 taps sit at rows 15..27
// of a zeroed column of 44, so tile row i of point p reads column row
// i - r_p + 15, always inside the column, with no per-element test (rows
// 28..31 of the 32 x 32 product read row 27's entries; they are not stored).
// SORTED: data[p] is the sample of point record p (the plan's sorted order,
// one coalesced load issued with the record's) instead of data[original index].
constexpr int kHalo = kBin + 2 * kK;          // 28 used rows / columns of a tile
constexpr int kTileCells = kHalo * kHalo;     // 784
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kSpCol = 44, kSpTap0 = 15, kSpGrp = 16;
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
struct SpreadWaveLds {
    uint2 a[kSpGrp][kSpCol];       // tap of d s w2: {re_hi, im_hi}, {re_lo, im_lo}
    unsigned b[kSpGrp][kSpCol];    // tap of w1: {hi, lo}
    unsigned rc[64];               // (15 - r_p) | (15 - c_p) << 16 of the wave's 64 points
};

__device__ __forceinline__ unsigned pack_h2(_Float16 lo16, _Float16 hi16)
{
    return (unsigned)__builtin_bit_cast(unsigned short, lo16) | ((unsigned)__builtin_bit_cast(unsigned short, hi16) << 16);
}

__device__ __forceinline__ void split_h(float x, _Float16& hi, _Float16& lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// A block spreads up to kSpUnit consecutive subproblems of one bin (up to
// kSpUnit * maxpts points, 256 per pass) into one tile, stored under the first
// of them: blocks s with (s - first subproblem of the bin) % kSpUnit != 0
// return at once, and k_spread_merge reads every kSpUnit-th tile of a bin.
// (A bin just over maxpts points then costs one block, not a second, nearly
// empty one.)
constexpr int kSpUnit = 2;
// Subproblem s (entry d = (bin, first point, count)) of a plan whose bins'
// first subproblems are subbeg and point totals tot; its tile goes to
// scratch[s].
template <bool SORTED>
__device__ __forceinline__ void spread_halo_body(int s, int4 d, const PtRec<float>* __restrict__ pts,
                                                 const float2* __restrict__ data, const int* __restrict__ subbeg,
                                                 const int* __restrict__ tot, int maxpts, const BinGeom& geo,
                                                 const WinArgs& wa, float2* __restrict__ scratch)
{
    __shared__ __attribute__((aligned(16))) union {
        SpreadWaveLds w[4];
        float red[4 * 32 * 32 * 2];   // [w][cell(32x32)][re, im], after the products
    } sm;
    const int k = s - subbeg[d.x];                          // index of subproblem s in its bin
    if (k % kSpUnit != 0) return;
    const int cnt = min(kSpUnit * maxpts, tot[d.x] - k * maxpts);   // points of the unit from d.y on
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // a unit of at most 64 points is wave 0's alone: the other waves leave
    // before any block barrier, and wave 0 stores its tile from its
    // accumulators (the 4-wave sum would add their zero tiles to it); sparse
    // point sets have ~16 points per bin
    const bool solo = cnt <= 64;
    if (solo && w != 0) return;
    const int O1 = (d.x % geo.nb1) * kBin, O2 = (d.x / geo.nb1) * kBin;
    SpreadWaveLds& L = sm.w[w];
    {
        constexpr int n16 = (int)((sizeof(L.a) + sizeof(L.b)) / 16);
        u4v* z = reinterpret_cast<u4v*>(&L);
        for (int e = lane; e < n16; e += 64) z[e] = u4v{0u, 0u, 0u, 0u};
    }
    // this lane's point of every pass (zeros beyond the unit), loads in flight together
    constexpr int kPass = (kSpUnit * kGatherThreads + 255) / 256;
    PtRec<float> rec[kPass];
    float2 v[kPass];
#pragma unroll
    for (int ps = 0; ps < kPass; ++ps) {
        const int q = 256 * ps + 64 * w + lane;
        rec[ps].t1 = rec[ps].t2 = 0.f;
        rec[ps].cell = rec[ps].idx = 0;
        v[ps] = make_float2(0.f, 0.f);
        if (q < cnt) {
            rec[ps] = pts[d.y + q];
            v[ps] = data[SORTED ? d.y + q : rec[ps].idx];
        }
    }
    // the wave's scale 2^(14 - e), e = exponent of its largest data component
    float mx = 0.f;
#pragma unroll
    for (int ps = 0; ps < kPass; ++ps) mx = fmaxf(mx, fmaxf(fabsf(v[ps].x), fabsf(v[ps].y)));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const int ex = min(max((int)((__float_as_uint(mx) >> 23) & 0xff) - 127, -100), 100);
    const float sc = __uint_as_float((unsigned)(127 + 14 - ex) << 23);
    const float inv_sc = __uint_as_float((unsigned)(127 - 14 + ex) << 23);
    f16v acc_re = {}, acc_im = {};
    const int i = lane & 31, kg = lane >> 5, ic = min(i, kHalo - 1);
#pragma unroll
    for (int ps = 0; ps < kPass; ++ps) {
        const int nw = min(64, max(0, cnt - 256 * ps - 64 * w));     // points of this wave in this pass
        if (nw == 0) continue;
        float w1[kW], w2[kW];
        int rl = 0, cl = 0;
        if (64 * w + lane < cnt - 256 * ps) {
            taps(rec[ps].t1, wa, w1);
            taps(rec[ps].t2, wa, w2);
            rl = min(max((rec[ps].cell >> 16) - O2, 0), kBin - 1);
            cl = min(max((rec[ps].cell & 0xffff) - O1, 0), kBin - 1);
        } else {
#pragma unroll
            for (int m = 0; m < kW; ++m) w1[m] = w2[m] = 0.f;
        }
        const float vx = v[ps].x * sc, vy = v[ps].y * sc;
        uint2 aw[kW];
        unsigned bw[kW];
#pragma unroll
        for (int m = 0; m < kW; ++m) {
            _Float16 rh, rlo, ih, ilo, bh, blo;
            split_h(vx * w2[m], rh, rlo);
            split_h(vy * w2[m], ih, ilo);
            split_h(w1[m], bh, blo);
            aw[m] = make_uint2(pack_h2(rh, ih), pack_h2(rlo, ilo));
            bw[m] = pack_h2(bh, blo);
        }
        L.rc[lane] = (unsigned)(kSpTap0 - rl) | ((unsigned)(kSpTap0 - cl) << 16);
        wave_sync_lds();
        for (int g = 0; g < nw; g += kSpGrp) {
            if ((lane >> 4) == (g >> 4)) {
                const int pl = lane & (kSpGrp - 1);
#pragma unroll
                for (int m = 0; m < kW; ++m) {
                    L.a[pl][kSpTap0 + m] = aw[m];
                    L.b[pl][kSpTap0 + m] = bw[m];
                }
            }
            wave_sync_lds();
            u4v ahr, ahi, alr, ali, bh, bl;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int p0 = 8 * kg + 2 * t, p1 = p0 + 1;
                const unsigned r0 = L.rc[g + p0], r1 = L.rc[g + p1];
                const uint2 a0 = L.a[p0][ic + (int)(r0 & 0xffffu)], a1 = L.a[p1][ic + (int)(r1 & 0xffffu)];
                const unsigned b0 = L.b[p0][ic + (int)(r0 >> 16)], b1 = L.b[p1][ic + (int)(r1 >> 16)];
                ahr[t] = __builtin_amdgcn_perm(a1.x, a0.x, 0x05040100u);
                ahi[t] = __builtin_amdgcn_perm(a1.x, a0.x, 0x07060302u);
                alr[t] = __builtin_amdgcn_perm(a1.y, a0.y, 0x05040100u);
                ali[t] = __builtin_amdgcn_perm(a1.y, a0.y, 0x07060302u);
                bh[t] = __builtin_amdgcn_perm(b1, b0, 0x05040100u);
                bl[t] = __builtin_amdgcn_perm(b1, b0, 0x07060302u);
            }
            const h8v Bh = __builtin_bit_cast(h8v, bh), Bl = __builtin_bit_cast(h8v, bl);
            acc_re = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, alr), Bh, acc_re, 0, 0, 0);
            acc_re = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, ahr), Bl, acc_re, 0, 0, 0);
            acc_re = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, ahr), Bh, acc_re, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, ali), Bh, acc_im, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, ahi), Bl, acc_im, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, ahi), Bh, acc_im, 0, 0, 0);
            wave_sync_lds();
        }
    }
    float2* out = scratch + (size_t)s * kTileCells;
    if (solo) {   // wave 0: unscale and store the 28x28 used part
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int ti = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), tj = lane & 31;
            if (ti < kHalo && tj < kHalo) out[ti * kHalo + tj] = make_float2(acc_re[r] * inv_sc, acc_im[r] * inv_sc);
        }
        return;
    }
    // unscale, reduce the 4 waves' tiles in a fixed order, store the 28x28 used part
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int ti = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), tj = lane & 31;
        sm.red[(w * 1024 + ti * 32 + tj) * 2] = acc_re[r] * inv_sc;
        sm.red[(w * 1024 + ti * 32 + tj) * 2 + 1] = acc_im[r] * inv_sc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTileCells; e += 256) {
        const int ti = e / kHalo, tj = e - ti * kHalo;
        const int c = ti * 32 + tj;
        float sr = sm.red[c * 2], si = sm.red[c * 2 + 1];
#pragma unroll
        for (int ww = 1; ww < 4; ++ww) {
            sr += sm.red[(ww * 1024 + c) * 2];
            si += sm.red[(ww * 1024 + c) * 2 + 1];
        }
        out[e] = make_float2(sr, si);
    }
}
template <bool SORTED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_spread_halo(
    const PtRec<float>* __restrict__ pts, const float2* __restrict__ data, const int4* __restrict__ sub,
    const int* __restrict__ nsub, const int* __restrict__ subbeg, const int* __restrict__ tot, int maxpts, BinGeom geo,
    WinArgs wa, float2* __restrict__ scratch)
{
    const int ns = *nsub;
    const int s = xcd_runs(blockIdx.x);
    if (s >= ns) return;
    spread_halo_body<SORTED>(s, sub[s], pts, data, subbeg, tot, maxpts, geo, wa, scratch);
}

// The complex128 form of k_spread_halo: the same tile
//     T[i][j] = sum_p (d_p w2_p[i - r_p]) * w1_p[j - c_p],
// the same units of up to kSpUnit subproblems of one bin and the same tile
// store, from v_mfma_f64_16x16x4f64 products (4 points per instruction, the
// 32 x 32 tile as 2 x 2 blocks of 16 x 16, real and imaginary parts apart;
// FP64 operands need no split and no scaling).  A wave stages its 64 points'
// 13 taps of d w2 and of w1 in LDS (tap-major, point fastest); in a 4-point
// step lane l takes point 4 g + l / 16 and row / column 16 h + l % 16 of block
// h, reading tap (row - r_p) of that point, or zero outside its window.
// Accumulator element r of lane l holds row 4 r + l / 16, column l % 16 of
// its block (measured with a known product).  The 4 waves' tiles are summed
// in a fixed order, as in k_spread_halo.
typedef double d4v __attribute__((ext_vector_type(4)));
struct SpreadWaveLds64 {
    double2 a[kW][64];    // d_p w2_p[m]
    double b[kW][64];     // w1_p[m]
    int r[64], c[64];     // r_p, c_p: the point's cell in the bin
};

template <bool SORTED>
__global__ __launch_bounds__(256) void k_spread_halo64(const PtRec<double>* __restrict__ pts,
                                                       const double2* __restrict__ data, const int4* __restrict__ sub,
                                                       const int* __restrict__ nsub, const int* __restrict__ subbeg,
                                                       const int* __restrict__ tot, int maxpts, BinGeom geo,
                                                       WinArgs wa, double2* __restrict__ scratch)
{
    __shared__ __attribute__((aligned(16))) union {
        SpreadWaveLds64 w[4];
        double red[4 * 32 * 32 * 2];   // [w][cell(32x32)][re, im], after the products
    } sm;
    const int s = blockIdx.x;
    if (s >= *nsub) return;
    const int4 d = sub[s];
    const int k = s - subbeg[d.x];                          // index of subproblem s in its bin
    if (k % kSpUnit != 0) return;
    const int cnt = min(kSpUnit * maxpts, tot[d.x] - k * maxpts);   // points of the unit from d.y on
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int O1 = (d.x % geo.nb1) * kBin, O2 = (d.x / geo.nb1) * kBin;
    SpreadWaveLds64& L = sm.w[w];
    const int col = lane & 15, q4 = lane >> 4;
    d4v acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc_re[bi][bj] = acc_im[bi][bj] = d4v{0.0, 0.0, 0.0, 0.0};
    for (int base = 64 * w; base < cnt; base += 256) {
        const int nw = min(64, cnt - base);                 // points of this wave in this pass
        double w1[kW], w2[kW];
        double2 v = make_double2(0.0, 0.0);
        int rl = 0, cl = 0;
        if (lane < nw) {
            const PtRec<double> rec = pts[d.y + base + lane];
            v = data[SORTED ? d.y + base + lane : rec.idx];
            taps(rec.t1, wa, w1);
            taps(rec.t2, wa, w2);
            rl = min(max((rec.cell >> 16) - O2, 0), kBin - 1);
            cl = min(max((rec.cell & 0xffff) - O1, 0), kBin - 1);
        } else {
#pragma unroll
            for (int m = 0; m < kW; ++m) w1[m] = w2[m] = 0.0;
        }
#pragma unroll
        for (int m = 0; m < kW; ++m) {
            L.a[m][lane] = make_double2(v.x * w2[m], v.y * w2[m]);
            L.b[m][lane] = w1[m];
        }
        L.r[lane] = rl;
        L.c[lane] = cl;
        wave_sync_lds();
        for (int g = 0; g < nw; g += 4) {
            const int p = g + q4;
            const int rp = L.r[p], cp = L.c[p];
            double ar[2], ai[2], bb[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int mi = 16 * h + col - rp, mj = 16 * h + col - cp;
                const bool ina = (unsigned)mi < (unsigned)kW, inb = (unsigned)mj < (unsigned)kW;
                const double2 av = L.a[ina ? mi : 0][p];
                const double bv = L.b[inb ? mj : 0][p];
                ar[h] = ina ? av.x : 0.0;
                ai[h] = ina ? av.y : 0.0;
                bb[h] = inb ? bv : 0.0;
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    acc_re[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[bi], bb[bj], acc_re[bi][bj], 0, 0, 0);
                    acc_im[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[bi], bb[bj], acc_im[bi][bj], 0, 0, 0);
                }
        }
        wave_sync_lds();                                    // staging reused by the next pass
    }
    // reduce the 4 waves' tiles in a fixed order and store the 28x28 used part
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ti = 16 * bi + 4 * r + q4, tj = 16 * bj + col;
                sm.red[(w * 1024 + ti * 32 + tj) * 2] = acc_re[bi][bj][r];
                sm.red[(w * 1024 + ti * 32 + tj) * 2 + 1] = acc_im[bi][bj][r];
            }
    __syncthreads();
    double2* out = scratch + (size_t)s * kTileCells;
    for (int e = threadIdx.x; e < kTileCells; e += 256) {
        const int ti = e / kHalo, tj = e - ti * kHalo;
        const int c = ti * 32 + tj;
        double sr = sm.red[c * 2], si = sm.red[c * 2 + 1];
#pragma unroll
        for (int ww = 1; ww < 4; ++ww) {
            sr += sm.red[(ww * 1024 + c) * 2];
            si += sm.red[(ww * 1024 + c) * 2 + 1];
        }
        out[e] = make_double2(sr, si);
    }
}

// Grid cell (n2, n1) = sum over the neighbouring bins b' (fixed order) and
// their tiles (every kSpUnit-th subproblem, in order) of
// tile_s[(n2 - O2' + K) mod L2][(n1 - O1' + K) mod L1] where that index lies
// in [0, 28) (merge_cell).  k_spread_merge writes every cell to the grid; a
// batch's matrix-core column pass sums the cells of its lines itself.
// The neighbours are the bins at offsets -R..R per dimension.  R = 1 covers
// every tile that reaches a bin when each bin is at least K cells wide; when
// the last bin of a dimension is thinner than K (L mod 16 = 2 or 4), the
// tiles of the bins on either side of it reach two bins across the wrap, and
// that dimension needs R = 2 (merge_radius).  With 2R + 1 > nb the offsets
// are 0..nb-1 instead, so that no bin is visited twice.
__host__ __device__ constexpr int merge_radius(int L) { return (L % kBin != 0 && L % kBin < kK) ? 2 : 1; }

template <typename T, int R1, int R2>
__device__ __forceinline__ cplx<T> merge_cell(int n2, int n1, const cplx<T>* __restrict__ scratch,
                                              const int* __restrict__ subbeg, const BinGeom& geo, int nb2)
{
    using C = cplx<T>;
    constexpr int W1 = 2 * R1 + 1, W2 = 2 * R2 + 1, NN = W1 * W2;
    const int nb1 = geo.nb1;
    const int b1 = n1 / kBin, b2 = n2 / kBin;
    const int s1 = W1 > nb1 ? 0 : b1 - R1 + nb1, c1 = W1 > nb1 ? nb1 : W1;   // first offset, count
    const int s2 = W2 > nb2 ? 0 : b2 - R2 + nb2, c2 = W2 > nb2 ? nb2 : W2;
    int beg[NN], end[NN], off[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        const int j2 = j / W1, j1 = j % W1;
        const int y2 = (s2 + j2) % nb2, x1 = (s1 + j1) % nb1;
        int r = n2 - (y2 * kBin - kK), c = n1 - (x1 * kBin - kK);
        r = r < 0 ? r + geo.L2 : (r >= geo.L2 ? r - geo.L2 : r);
        c = c < 0 ? c + geo.L1 : (c >= geo.L1 ? c - geo.L1 : c);
        const int nb = y2 * nb1 + x1;
        const bool in = j1 < c1 && j2 < c2 && r < kHalo && c < kHalo;
        beg[j] = in ? subbeg[nb] : 0;
        end[j] = in ? subbeg[nb + 1] : 0;
        off[j] = r * kHalo + c;
    }
    // tiles are summed in (neighbour, subproblem) order; the loads go out in
    // groups of 8 so that a bin with many subproblems does not pay the memory
    // latency once per tile
    T sr = 0, si = 0;
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        int s = beg[j];
        for (; s + 7 * kSpUnit < end[j]; s += 8 * kSpUnit) {
            C v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = scratch[(size_t)(s + q * kSpUnit) * kTileCells + off[j]];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                sr += v[q].x;
                si += v[q].y;
            }
        }
        for (; s < end[j]; s += kSpUnit) {
            const C v = scratch[(size_t)s * kTileCells + off[j]];
            sr += v.x;
            si += v.y;
        }
    }
    return mk(sr, si);
}
// One block per bin writes the bin's cells.
template <typename T, int R1, int R2>
__device__ __forceinline__ void spread_merge_body(int bin, const cplx<T>* __restrict__ scratch,
                                                  const int* __restrict__ subbeg, const BinGeom& geo, int nb2,
                                                  cplx<T>* __restrict__ Gt)
{
    const int nb1 = geo.nb1;
    const int b1 = bin % nb1, b2 = bin / nb1;
    const int n2 = b2 * kBin + threadIdx.x / kBin, n1 = b1 * kBin + threadIdx.x % kBin;
    if (n2 >= geo.L2 || n1 >= geo.L1) return;
    Gt[(size_t)n2 * geo.L1 + n1] = merge_cell<T, R1, R2>(n2, n1, scratch, subbeg, geo, nb2);
}
template <typename T, int R1, int R2>
__global__ __launch_bounds__(256) void k_spread_merge(const cplx<T>* __restrict__ scratch,
                                                      const int* __restrict__ subbeg, BinGeom geo, int nb2,
                                                      cplx<T>* __restrict__ Gt)
{
    spread_merge_body<T, R1, R2>(xcd_block(blockIdx.x, gridDim.x), scratch, subbeg, geo, nb2, Gt);
}

}  // namespace nfft2d
