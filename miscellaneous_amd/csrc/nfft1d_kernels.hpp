// nfft1d_kernels.hpp -- gfx950 kernels of the batched 1-D NFFTs (complex64).
//
// A batch is B lines of N modes each.  The 1-D oversampled grid of a line has
// L = 2N cells; it is kept unshifted (spatial index n in [0, L), mode k at
// k mod L), as the 2-D grid is, so the row passes of the 2-D transform
// (k_t2_rows: pre-compensation, zero padding and the pruned 2N-point FFT;
// k_t1_rows: FFT, crop and deconvolution) serve the 1-D lines unchanged, with
// the line index as the row.  What is 1-D here is the window step:
//   k_gather1d  y_l = sum_{m=0}^{12} w_l[m] G[(mu_l + m - K) mod L]
//               (BP_FBP_FFBP/NFFT1_1D.m:47-48), the line staged in LDS;
//   k_spread1d  G[(mu_l + m - K) mod L] += d_l w_l[m]
//               (NFFT/NFFT2/1D/Matlab/NFFT2_1D.m:59-78), summed in LDS as
//               64-bit fixed point with integer atomics: exact, so the result
//               does not depend on the order of the additions.
// mu_l = rint(c x_l) (half to even, as the 2-D path; NFFT1_1D.m's round
// differs only at exact ties, either is a valid NFFT).
#pragma once

#include "nfft2d_kernels.hpp"

namespace nfft2d {

constexpr int k1dThreads = 256;
constexpr int k1dPointsPerBlock = 1024;   // points of one line per gather block

// grid: (lines, ceil(M / k1dPointsPerBlock)); dynamic LDS L complex.
__global__ __launch_bounds__(k1dThreads) void k_gather1d(const float2* __restrict__ G, int L, float invL,
                                                         const float* __restrict__ x, int M, WinArgs wa,
                                                         float2* __restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    float2* row = reinterpret_cast<float2*>(smem_raw);
    const size_t b = blockIdx.x;
    const float2* g = G + b * (size_t)L;
    for (int i = threadIdx.x; i < L; i += k1dThreads) row[i] = g[i];
    __syncthreads();
    const float* xb = x + b * (size_t)M;
    float2* ob = out + b * (size_t)M;
    const int l0 = blockIdx.y * k1dPointsPerBlock;
#pragma unroll
    for (int q = 0; q < k1dPointsPerBlock / k1dThreads; ++q) {
        const int l = l0 + q * k1dThreads + (int)threadIdx.x;
        if (l < M) {
            int mu;
            float t;
            locate(xb[l], L, invL, mu, t);
            float w[kW];
            taps(t, wa, w);
            float ax = 0.f, ay = 0.f;
#pragma unroll
            for (int m = 0; m < kW; ++m) {
                int i = mu + m - kK;                                 // L >= 16 > K: one wrap at most
                i += i < 0 ? L : (i >= L ? -L : 0);
                ax = fmaf(w[m], row[i].x, ax);
                ay = fmaf(w[m], row[i].y, ay);
            }
            ob[l] = make_float2(ax, ay);
        }
    }
}

// One block per line; dynamic LDS 2 L 64-bit accumulators.  The samples are
// scaled by 2^s, s chosen from the line's largest component and its point
// count so that no cell's sum can leave the signed 64-bit range (taps are
// below 1): every contribution rounds to an integer once and the sums are
// exact.
__global__ __launch_bounds__(k1dThreads) void k_spread1d(const float2* __restrict__ data, const float* __restrict__ x,
                                                         int M, int L, float invL, WinArgs wa, float2* __restrict__ G)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem_raw);   // [L][re, im]
    __shared__ float s_max[k1dThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t b = blockIdx.x;
    const float2* d = data + b * (size_t)M;
    const float* xb = x + b * (size_t)M;
    for (int i = tid; i < 2 * L; i += k1dThreads) acc[i] = 0ull;
    float mx = 0.f;
    for (int l = tid; l < M; l += k1dThreads) {
        const float2 v = d[l];
        mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) s_max[w] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    // |contribution| < 2^(ex + 1) with ex the largest exponent; M of them fit
    // below 2^62 after scaling by 2^(61 - ex - ceil(log2(M + 1)))
    const int ex = min(max((int)((__float_as_uint(mx) >> 23) & 0xff) - 127, -400), 400);
    const int lg = 32 - __clz(M);                                     // >= log2(M + 1)
    const int sh = 61 - ex - lg;
    const double scale = ldexp(1.0, sh), inv = ldexp(1.0, -sh);
    for (int l = tid; l < M; l += k1dThreads) {
        int mu;
        float t;
        locate(xb[l], L, invL, mu, t);
        float wt[kW];
        taps(t, wa, wt);
        const float2 v = d[l];
#pragma unroll
        for (int m = 0; m < kW; ++m) {
            int i = mu + m - kK;
            i += i < 0 ? L : (i >= L ? -L : 0);
            const long long re = __double2ll_rn((double)(v.x * wt[m]) * scale);
            const long long im = __double2ll_rn((double)(v.y * wt[m]) * scale);
            atomicAdd(&acc[2 * i], (unsigned long long)re);
            atomicAdd(&acc[2 * i + 1], (unsigned long long)im);
        }
    }
    __syncthreads();
    float2* g = G + b * (size_t)L;
    for (int i = tid; i < L; i += k1dThreads)
        g[i] = make_float2((float)((double)(long long)acc[2 * i] * inv), (float)((double)(long long)acc[2 * i + 1] * inv));
}

}  // namespace nfft2d
