// nfft2d_device.hpp -- device building blocks for the gfx950 2D NFFT kernels.
//
// Complex arithmetic on float2/double2, the Knab window (per-tap polynomials:
// degree 8 for FP32, 14 for FP64), and the register/LDS radix-2/4/8 Stockham
// butterflies used by the FFT passes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nfft2d {

constexpr int kK = 6;                 // window half-width (NFFT/NFFT2_2D.cu:21)
constexpr int kW = 2 * kK + 1;        // taps per dimension
constexpr int kC = 2;                 // oversampling factor (NFFT/NFFT2_2D.cu:20)
constexpr int kPolyDeg = 8;           // degree of the per-tap window polynomial
constexpr int kNE = kPolyDeg / 2 + 1; // even-part coefficients (in t^2)
constexpr int kNO = kPolyDeg / 2;     // odd-part coefficients (in t^2)

// ---------------------------------------------------------------- complex types
template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using cplx = typename cvec<T>::type;

__device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk(a.x, -a.y); }
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { return mk(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse): SIGN = -1 gives -i * a
template <int SIGN, typename C> __device__ __forceinline__ C mul_i(C a)
{
    return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// ---------------------------------------------------------------- window
// FP32 window: for tap offset j in [0, K], P_j(t) = phi_hat(t - j) / S for the
// point's fractional offset t = c*x - rint(c*x) in [-1/2, 1/2], written as
// E_j(t^2) + t O_j(t^2); tap -j is E_j(t^2) - t O_j(t^2) because phi_hat is
// even.  Coefficients are fitted on the host in FP64 (Chebyshev interpolation,
// degree 8, max error ~1e-9 of the peak).  S = I0(K alpha) keeps values O(1);
// the deconvolution tables carry the same factor.
struct WinPoly {
    float e[kK + 1][kNE];
    float o[kK + 1][kNO];
};

// w[m] for m = 0..2K is the weight of grid index mu + (m - K).
__device__ __forceinline__ void window_taps(float t, const WinPoly& wp, float (&w)[kW])
{
    const float u = t * t;
#pragma unroll
    for (int j = 0; j <= kK; ++j) {
        float e = wp.e[j][kNE - 1];
#pragma unroll
        for (int q = kNE - 2; q >= 0; --q) e = fmaf(e, u, wp.e[j][q]);
        if (j == 0) {
            w[kK] = e;
        } else {
            float o = wp.o[j][kNO - 1];
#pragma unroll
            for (int q = kNO - 2; q >= 0; --q) o = fmaf(o, u, wp.o[j][q]);
            w[kK + j] = fmaf(t, o, e);
            w[kK - j] = fmaf(-t, o, e);
        }
    }
}

// FP64 window: the same split, degree 14 (max error ~4e-15 of the peak, the
// rounding level of the FP64 closed form NFFT/NFFT2_2D.cu:139-149 it
// replaces), with the value and derivative tables in device memory (one
// 1680-byte copy per FP64 plan, read through uniform loads), so the FP32
// kernels' arguments do not grow.
constexpr int kPolyDeg64 = 14;
constexpr int kNE64 = kPolyDeg64 / 2 + 1;
constexpr int kNO64 = kPolyDeg64 / 2;
struct WinPoly64 {
    double e[kK + 1][kNE64];
    double o[kK + 1][kNO64];
};
struct WinTables64 {
    WinPoly64 val;    // phi_hat / S
    WinPoly64 der;    // phi_hat' / S
};

__device__ __forceinline__ void window_taps64(double t, const WinPoly64& wp, double (&w)[kW])
{
    const double u = t * t;
#pragma unroll
    for (int j = 0; j <= kK; ++j) {
        double e = wp.e[j][kNE64 - 1];
#pragma unroll
        for (int q = kNE64 - 2; q >= 0; --q) e = fma(e, u, wp.e[j][q]);
        if (j == 0) {
            w[kK] = e;
        } else {
            double o = wp.o[j][kNO64 - 1];
#pragma unroll
            for (int q = kNO64 - 2; q >= 0; --q) o = fma(o, u, wp.o[j][q]);
            w[kK + j] = fma(t, o, e);
            w[kK - j] = fma(-t, o, e);
        }
    }
}

// Derivative taps d/dt phi_hat(t - (m - K)) / S: tap -j is -E'_j(t^2) + t O'_j(t^2).
__device__ __forceinline__ void window_dtaps64(double t, const WinPoly64& wp, double (&w)[kW])
{
    const double u = t * t;
#pragma unroll
    for (int j = 0; j <= kK; ++j) {
        double o = wp.o[j][kNO64 - 1];
#pragma unroll
        for (int q = kNO64 - 2; q >= 0; --q) o = fma(o, u, wp.o[j][q]);
        if (j == 0) {
            w[kK] = t * o;
        } else {
            double e = wp.e[j][kNE64 - 1];
#pragma unroll
            for (int q = kNE64 - 2; q >= 0; --q) e = fma(e, u, wp.e[j][q]);
            w[kK + j] = fma(t, o, e);
            w[kK - j] = fma(t, o, -e);
        }
    }
}

// ---------------------------------------------------------------- point -> (mu, t)
// mu = rint(c x) (half to even, NFFT/NFFT2_2D.cu:168), t = c x - mu in
// [-1/2, 1/2] exactly, and mu wrapped into [0, L) (the reference's positive
// `modulo`, :124-125).  The wrap uses a floating-point quotient estimate and an
// exact integer fix-up (no integer division).  Non-finite or huge coordinates
// map to cell 0 with t = NaN so the result is NaN and every index stays in
// bounds.
template <typename T>
__device__ __forceinline__ void locate(T x, int L, T invL, int& mu_wrapped, T& t)
{
    T cx = (T)kC * x;
    if (!(fabs(cx) < (T)1073741824)) {   // also catches NaN
        mu_wrapped = 0;
        t = (T)NAN;
        return;
    }
    const T mu = rint(cx);
    t = cx - mu;
    const int mi = (int)mu;
    int r = mi - (int)floor(mu * invL) * L;
    while (r < 0) r += L;      // the estimate is exact or one off except for |c x| >~ 2^24
    while (r >= L) r -= L;
    mu_wrapped = r;
}

// ---------------------------------------------------------------- butterflies
template <int SIGN, typename C>
__device__ __forceinline__ void dft2(C& a, C& b)
{
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int SIGN, typename C>
__device__ __forceinline__ void dft4(C& a0, C& a1, C& a2, C& a3)
{
    C s02 = cadd(a0, a2), d02 = csub(a0, a2);
    C s13 = cadd(a1, a3), d13 = mul_i<SIGN>(csub(a1, a3));
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
}

template <int SIGN, typename C>
__device__ __forceinline__ void dft8(C (&v)[8])
{
    using T = decltype(v[0].x);
    const T h = (T)0.70710678118654752440084436210485;
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    // twiddles W8^k, W8 = e^{SIGN i pi/4}
    C t1 = mk((o1.x - SIGN * o1.y) * h, (o1.y + SIGN * o1.x) * h);   // o1 * (1 + SIGN i)/sqrt2
    C t2 = mul_i<SIGN>(o2);
    C t3 = mk((-o3.x - SIGN * o3.y) * h, (-o3.y + SIGN * o3.x) * h);  // o3 * (-1 + SIGN i)/sqrt2
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1); v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3); v[7] = csub(e3, t3);
}

template <int R, int SIGN, typename C>
__device__ __forceinline__ void dftR(C* v)
{
    if constexpr (R == 2) {
        dft2<SIGN>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<SIGN>(v[0], v[1], v[2], v[3]);
    } else {
        C (&a)[8] = *reinterpret_cast<C(*)[8]>(v);
        dft8<SIGN>(a);
    }
}

// ---------------------------------------------------------------- Stockham plan
// N = 2^n points, 8 elements per thread per line: floor(n/3) radix-8 stages,
// preceded by one radix-2 or radix-4 stage when n mod 3 != 0.
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
template <int N> struct FftPlan {
    static constexpr int n = ilog2(N);
    static constexpr int r0 = (n % 3 == 0) ? 8 : (n % 3 == 1 ? 2 : 4);
    static constexpr int stages = (n % 3 == 0) ? n / 3 : n / 3 + 1;
    static constexpr int tpl = N / 8;        // threads per line
    static constexpr int radix(int s) { return s == 0 ? r0 : 8; }
    // Ns before stage s (product of the radices of the earlier stages)
    static constexpr int ns(int s) { return s == 0 ? 1 : r0 * (1 << (3 * (s - 1))); }
};

// LDS padding: one extra element every 16 breaks power-of-two strides.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
template <int N> constexpr int lds_line() { return N + N / 16; }

// Twiddle table of size TWN: tw[k] = e^{-2 pi i k / TWN} (FP64-accurate,
// rounded to T).  The inverse direction conjugates.
template <int SIGN, typename C>
__device__ __forceinline__ C twid(const C* __restrict__ tw, int idx)
{
    C w = tw[idx];
    return SIGN < 0 ? w : cconj(w);
}

// One Stockham stage on V lines that share the butterfly positions.  Inputs
// are v[line][b*R + r] = x[bi + r*N/R] with bi = j + b*(N/8); on return
// v[line][b*R + r] holds the stage output destined for index
// (bi/Ns)*Ns*R + (bi%Ns) + r*Ns.  `tw` has TWN entries, TWN a multiple of N.
template <int N, int R, int NS, int V, int SIGN, typename C>
__device__ __forceinline__ void stage_compute(C (&v)[V][8], int j, const C* __restrict__ tw, int TWN)
{
    constexpr int B = 8 / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int bi = j + b * (N / 8);
        const int k = bi % NS;
        if constexpr (NS > 1) {
            const int step = TWN / (NS * R);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = twid<SIGN>(tw, (r * k * step) % TWN);
#pragma unroll
                for (int l = 0; l < V; ++l) v[l][b * R + r] = cmul(v[l][b * R + r], w);
            }
        }
#pragma unroll
        for (int l = 0; l < V; ++l) dftR<R, SIGN>(&v[l][b * R]);
    }
}

template <int N, int R, int NS>
__device__ __forceinline__ int stage_out_index(int j, int b, int r)
{
    const int bi = j + b * (N / 8);
    return (bi / NS) * NS * R + (bi % NS) + r * NS;
}

template <int N, int R>
__device__ __forceinline__ int stage_in_index(int j, int b, int r)
{
    return j + b * (N / 8) + r * (N / R);
}

// Middle stages (1 .. stages-2) run entirely through LDS: read, compute, write.
// `lines` points at V consecutive lines of lds_line<N>() elements each.
// Caller is responsible for a barrier before the first call.
template <int N, int S, int V, int SIGN, typename C>
__device__ __forceinline__ void stage_lds(C* lines, int j, const C* __restrict__ tw, int TWN)
{
    using P = FftPlan<N>;
    constexpr int R = P::radix(S);
    constexpr int NS = P::ns(S);
    constexpr int B = 8 / R;
    constexpr int LL = lds_line<N>();
    C v[V][8];
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int l = 0; l < V; ++l) v[l][b * R + r] = lines[l * LL + lpad(stage_in_index<N, R>(j, b, r))];
    __syncthreads();
    stage_compute<N, R, NS, V, SIGN>(v, j, tw, TWN);
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int l = 0; l < V; ++l) lines[l * LL + lpad(stage_out_index<N, R, NS>(j, b, r))] = v[l][b * R + r];
    __syncthreads();
}

template <int N, int S, int V, int SIGN, typename C>
__device__ __forceinline__ void stages_mid(C* lines, int j, const C* __restrict__ tw, int TWN)
{
    if constexpr (S < FftPlan<N>::stages - 1) {
        stage_lds<N, S, V, SIGN>(lines, j, tw, TWN);
        stages_mid<N, S + 1, V, SIGN>(lines, j, tw, TWN);
    }
}

// Full FFT of V lines whose inputs the caller has placed in registers in the
// first-stage order (v[l][b*R0 + r] = x[j + b*N/8 + r*N/R0]).  Runs the first
// stage in registers, the middle stages through LDS, and the last stage's
// butterflies in registers: on return v[l][b*RL + r] = X[j + b*N/8 + r*N/RL]
// (natural order), RL = radix of the last stage.
// A thread may serve a different line group and thread-in-line index in the
// first stage (lds_first, j_first) than in the later ones (lds_last, j_last):
// the LDS exchange in between makes the remap free, which is how the column
// passes load and store along different axes.  Every group's V lines are
// lds_line<N>() elements apart.  Ends without a barrier: the caller must sync
// before reusing the LDS.
template <int N, int V, int SIGN, typename C>
__device__ __forceinline__ void fft_lines(C (&v)[V][8], C* lds_first, int j_first, C* lds_last, int j_last,
                                          const C* __restrict__ tw, int TWN)
{
    using P = FftPlan<N>;
    constexpr int LL = lds_line<N>();
    constexpr int R0 = P::radix(0);
    stage_compute<N, R0, 1, V, SIGN>(v, j_first, tw, TWN);
    if constexpr (P::stages > 1) {
        constexpr int B0 = 8 / R0;
#pragma unroll
        for (int b = 0; b < B0; ++b)
#pragma unroll
            for (int r = 0; r < R0; ++r)
#pragma unroll
                for (int l = 0; l < V; ++l)
                    lds_first[l * LL + lpad(stage_out_index<N, R0, 1>(j_first, b, r))] = v[l][b * R0 + r];
        __syncthreads();
        stages_mid<N, 1, V, SIGN>(lds_last, j_last, tw, TWN);
        constexpr int SL = P::stages - 1;
        constexpr int RL = P::radix(SL);
        constexpr int NSL = P::ns(SL);
        constexpr int BL = 8 / RL;
#pragma unroll
        for (int b = 0; b < BL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r)
#pragma unroll
                for (int l = 0; l < V; ++l)
                    v[l][b * RL + r] = lds_last[l * LL + lpad(stage_in_index<N, RL>(j_last, b, r))];
        stage_compute<N, RL, NSL, V, SIGN>(v, j_last, tw, TWN);
    }
}

// Index of element (b, r) in the first-stage input order / last-stage output order.
template <int N>
__device__ __forceinline__ int first_in_index(int j, int b, int r)
{
    constexpr int R0 = FftPlan<N>::radix(0);
    return j + b * (N / 8) + r * (N / R0);
}
template <int N>
__device__ __forceinline__ int last_out_index(int j, int b, int r)
{
    using P = FftPlan<N>;
    constexpr int RL = P::radix(P::stages - 1);
    return j + b * (N / 8) + r * (N / RL);
}

}  // namespace nfft2d
