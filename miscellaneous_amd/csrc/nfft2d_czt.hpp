// nfft2d_czt.hpp -- the four FFT passes of a plan for line lengths that are
// not powers of two, as chirp-z (Bluestein) transforms.
//
// Every pass maps P inputs of a line to Q outputs
//     Y[q] = sum_{p<P} x[p] w^{(p + p0)(q + q0)},  w = e^{-2 pi i / L},  L = 2N,
// (type-2 rows / columns: P = N, Q = L, p0 = -N/2, q0 = 0, modes centred;
//  type-1 columns / rows: P = L, Q = N, p0 = 0, q0 = -N/2), which with
// ab = (a^2 + b^2 - (b - a)^2) / 2 and c[k] = w^{k^2/2} = e^{-i pi k^2 / L} is
//     Y[q] = c[q + q0] sum_p (x[p] c[p + p0]) conj(c[(q - p) + q0 - p0]),
// a linear convolution, taken circularly at the power of two m >= P + Q - 1
// through two LDS-resident FFTs of m points.  The inverse direction (e^{+})
// conjugates the input and the output.  The chirps (P and Q entries), the
// transform of the convolution kernel (m entries) and the FFT's twiddles
// (e^{-2 pi i t / m} = T1[t / 64] T2[t % 64]) are computed on the host in FP64
// at plan creation.  One block per line; O(m log m) per line instead of the
// O(N L) of the direct passes.
#pragma once

#include "nfft2d_device.hpp"

namespace nfft2d {

template <typename T>
struct CztTables {
    const cplx<T> *cin = nullptr;    // c[p + p0], P
    const cplx<T> *cout = nullptr;   // c[q + q0], Q
    const cplx<T> *hhat = nullptr;   // FFT_m of the kernel conj(c[d + q0 - p0]), d = q - p
    const cplx<T> *t1 = nullptr;     // e^{-2 pi i 64 j / m}, max(1, m / 64)
    const cplx<T> *t2 = nullptr;     // e^{-2 pi i j / m}, 64
    int P = 0, Q = 0, m = 0;
};

__host__ __device__ constexpr int czt_lds_entries(int m) { return m + (m >= 64 ? m / 64 : 1) + 64; }

// Smallest power of two m >= 3N - 1 (P + Q - 1 = N + 2N - 1): the chirp-z FFT length of a line of N.
__host__ __device__ constexpr int czt_m(int N)
{
    int m = 1;
    while (m < 3 * N - 1) m <<= 1;
    return m;
}

// The chirp-z tables of a line length N inside one table buffer (layout
// [cin2 N][cout2 L][hhat2 m][cin1 L][cout1 N][hhat1 m][T1 m/64][T2 64],
// L = 2N), for its type-2 (t2 = true) or type-1 pass.
template <typename T>
__host__ __device__ inline CztTables<T> czt_tables_of(const void* buf, int N, int m, bool t2)
{
    using C = cplx<T>;
    const int L = 2 * N;
    const C* b = (const C*)buf;
    CztTables<T> t;
    t.m = m;
    const C* tw = b + 3 * L + 2 * m;   // after both passes' chirps and kernels
    t.t1 = tw;
    t.t2 = tw + (m >= 64 ? m / 64 : 1);
    if (t2) {
        t.P = N;
        t.Q = L;
        t.cin = b;
        t.cout = b + N;
        t.hhat = b + N + L;
    } else {
        const C* o = b + N + L + m;
        t.P = L;
        t.Q = N;
        t.cin = o;
        t.cout = o + L;
        t.hhat = o + L + N;
    }
    return t;
}

// In-place radix-4 (+ one radix-2) Stockham FFT of m = 2^a points in LDS
// (Govindaraju et al., SC 2008): stage with span Ns reads v_r = X[j + r m/R],
// twiddles it by e^{-2 pi i r (j mod Ns) / (Ns R)}, and writes the R-point DFT
// to X[(j - j mod Ns) R + j mod Ns + r Ns]; all reads of a stage precede its
// writes (barrier).  Each thread holds at most 4 U values (blockDim >= m / (4 U)):
// U = 2 up to m = 8192 at 1024 threads, U = 4 for m = 16384.
template <typename T, int U>
__device__ void czt_fft(cplx<T> *X, int m, const cplx<T> *t1, const cplx<T> *t2)
{
    using C = cplx<T>;
    const int NT = blockDim.x, tid = threadIdx.x;
    auto tw = [&](int t) { return cmul(t1[t >> 6], t2[t & 63]); };
    for (int Ns = 1; Ns < m;) {
        const int R = (m / Ns) >= 4 ? 4 : 2;
        const int nb = m / R, per = (nb + NT - 1) / NT;
        const int tstep = m / (Ns * R);
        C v[4 * U];
        if (R == 4) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = tid + u * NT;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1);
                    C a0 = X[j], a1 = X[j + nb], a2 = X[j + 2 * nb], a3 = X[j + 3 * nb];
                    if (Ns > 1) {
                        a1 = cmul(a1, tw(k * tstep));
                        a2 = cmul(a2, tw(2 * k * tstep));
                        a3 = cmul(a3, tw(3 * k * tstep));
                    }
                    const C b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3);
                    const C d = csub(a1, a3);
                    const C b3 = mk(d.y, -d.x);   // (a1 - a3) (-i)
                    v[4 * u + 0] = cadd(b0, b2);
                    v[4 * u + 1] = cadd(b1, b3);
                    v[4 * u + 2] = csub(b0, b2);
                    v[4 * u + 3] = csub(b1, b3);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = tid + u * NT;
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
            for (int u = 0; u < 2 * U; ++u) {
                const int j = tid + u * NT;
                if (u < per && j < nb) {
                    const int k = j & (Ns - 1);
                    C a0 = X[j], a1 = X[j + nb];
                    if (Ns > 1) a1 = cmul(a1, tw(k * tstep));
                    v[2 * u] = cadd(a0, a1);
                    v[2 * u + 1] = csub(a0, a1);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 2 * U; ++u) {
                const int j = tid + u * NT;
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

// Line b of a chirp-z pass, by the whole block (smem: czt_lds_entries(m)
// entries; blockDim.x >= m / (4 U)).  The line reads x[p] = in[b * in_ls +
// p * in_es] and writes out[b * out_ls + q * out_es].  Type-2 rows scale the
// input by sl[b] se[p] (the pre-compensation D1[k1] D2[k2]); type-1 rows
// scale the output by sl[b] se[q] (the deconvolution); null: no scaling.
// Every butterfly and product is the same whatever the block size, so plans
// and batches (other block sizes) get the same bits.
template <typename T, int SIGN, bool SCALE_IN, int U>
__device__ __forceinline__ void czt_line(int b, cplx<T>* smem, const cplx<T>* __restrict__ in, long long in_ls,
                                         long long in_es, cplx<T>* __restrict__ out, long long out_ls,
                                         long long out_es, const T* __restrict__ sl, const T* __restrict__ se,
                                         const CztTables<T>& tb)
{
    using C = cplx<T>;
    C *X = smem;
    C *t1 = X + tb.m, *t2 = t1 + (tb.m >= 64 ? tb.m / 64 : 1);
    const int NT = blockDim.x, m = tb.m;
    for (int k = threadIdx.x; k < 64; k += NT) t2[k] = tb.t2[k];
    for (int k = threadIdx.x; k < (m >= 64 ? m / 64 : 1); k += NT) t1[k] = tb.t1[k];
    const T lsc = sl ? sl[b] : (T)1;
    for (int p = threadIdx.x; p < m; p += NT) {
        C v = mk((T)0, (T)0);
        if (p < tb.P) {
            v = in[b * in_ls + p * in_es];
            if (SCALE_IN && sl) v = cscale(v, lsc * se[p]);
            if (SIGN > 0) v = cconj(v);
            v = cmul(v, tb.cin[p]);
        }
        X[p] = v;
    }
    __syncthreads();
    czt_fft<T, U>(X, m, t1, t2);
    for (int k = threadIdx.x; k < m; k += NT) X[k] = cconj(cmul(X[k], tb.hhat[k]));
    __syncthreads();
    czt_fft<T, U>(X, m, t1, t2);
    const T inv = (T)1 / (T)m;
    for (int q = threadIdx.x; q < tb.Q; q += NT) {
        C v = cmul(cscale(cconj(X[q]), inv), tb.cout[q]);
        if (SIGN > 0) v = cconj(v);
        if (!SCALE_IN && sl) v = cscale(v, lsc * se[q]);
        out[b * out_ls + q * out_es] = v;
    }
}

// One line per block.
template <typename T, int SIGN, bool SCALE_IN, int U>
__global__ __launch_bounds__(1024) void k_czt(const cplx<T> *__restrict__ in, long long in_ls, long long in_es,
                                              cplx<T> *__restrict__ out, long long out_ls, long long out_es,
                                              const T *__restrict__ sl, const T *__restrict__ se, CztTables<T> tb)
{
    extern __shared__ unsigned char smem_raw[];
    czt_line<T, SIGN, SCALE_IN, U>((int)blockIdx.x, reinterpret_cast<cplx<T> *>(smem_raw), in, in_ls, in_es, out,
                                   out_ls, out_es, sl, se, tb);
}

}  // namespace nfft2d
