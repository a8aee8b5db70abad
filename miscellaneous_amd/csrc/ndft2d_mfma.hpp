// ndft2d_mfma.hpp -- type-2 NDFT (modes -> points) as a separable contraction on MFMA.
//
//   y_m = sum_{k1,k2} F[k1][k2] e^{-2 pi i (k1 x1_m / N1 + k2 x2_m / N2)}
//       = sum_{k1} A[m][k1] T[m][k1],   T = B F^T,
//   A[m][k1] = e^{-2 pi i k1 x1_m / N1},  B[m][k2] = e^{-2 pi i k2 x2_m / N2},
//
// the (M x N2)(N2 x N1) o (M x N1) form of the reference's direct sum
// (NFFT/NDFT2_2D.cu:15-32 fills the M x N kernel matrix that cublasZgemv
// contracts at :67; NFFT/NFFT2/Matlab/NDFT2_2D.m:9-14 is the same sum).
//
// T is one real GEMM: rows = points, K = 2 N2 (Re/Im of B interleaved per k2),
// columns = 2 N1 (the Re and Im parts of T for every k1).  Its left operand
// (the points' phasors) is generated in registers; its right operand (the
// modes, G below) is prepared once per call in the exact per-lane order the
// MFMA reads it.  The k1 contraction with A happens in the epilogue.
//
// Precisions (the operand type of the T product; accumulation is FP32):
//   kNdftFp32   v_mfma_f32_32x32x2f32, operands FP32;
//   kNdftBf16x3 v_mfma_f32_32x32x16_bf16, x = hi + lo split of both operands,
//               T = hi*hi + hi*lo + lo*hi;
//   kNdftFp16   v_mfma_f32_32x32x16_f16, each k1 row of F scaled by a power of
//               two into the FP16 range, the scale undone in the epilogue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ndft {

enum { kNdftFp32 = 0, kNdftBf16x3 = 1, kNdftFp16 = 2 };

constexpr int kWaves = 4;                  // waves per workgroup

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// One lane's slice of one 32-column x 16-K tile of the right operand.
template <int P> struct Slice;
template <> struct Slice<kNdftFp32> { float v[8]; };
template <> struct __attribute__((aligned(16))) Slice<kNdftBf16x3> { bf8v hi, lo; };
template <> struct __attribute__((aligned(16))) Slice<kNdftFp16> { h8v h; };

// Left-operand registers of one K-step.
template <int P> struct Left;
template <> struct Left<kNdftFp32> { float v[8]; };
template <> struct Left<kNdftBf16x3> { bf8v hi, lo; };
template <> struct Left<kNdftFp16> { h8v h; };

__device__ __forceinline__ float bf16_to_f32(__bf16 b)
{
    return __uint_as_float((uint32_t)__builtin_bit_cast(unsigned short, b) << 16);
}

// x = hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x - hi).
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - bf16_to_f32(hi));
}

template <int P> __device__ __forceinline__ void make_left(const float (&v)[8], Left<P>& o);
template <> __device__ __forceinline__ void make_left<kNdftFp32>(const float (&v)[8], Left<kNdftFp32>& o)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) o.v[e] = v[e];
}
template <> __device__ __forceinline__ void make_left<kNdftBf16x3>(const float (&v)[8], Left<kNdftBf16x3>& o)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        __bf16 h, l;
        split_bf16(v[e], h, l);
        o.hi[e] = h;
        o.lo[e] = l;
    }
}
template <> __device__ __forceinline__ void make_left<kNdftFp16>(const float (&v)[8], Left<kNdftFp16>& o)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) o.h[e] = (_Float16)v[e];
}

// acc += L * R over one 16-deep K-step, as NPROD<P> MFMA products (mma_part
// k = 0 .. NPROD - 1).  The K order inside the step is the same for both
// operands (lane group h, element e), so the sum is the full one.
template <int P> constexpr int nprod() { return P == kNdftFp32 ? 8 : (P == kNdftBf16x3 ? 3 : 1); }
template <int P, int K> __device__ __forceinline__ void mma_part(const Left<P>& a, const Slice<P>& b, f16v& acc)
{
    if constexpr (P == kNdftFp32) {
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[K], b.v[K], acc, 0, 0, 0);
    } else if constexpr (P == kNdftBf16x3) {
        if constexpr (K == 0) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.lo, b.hi, acc, 0, 0, 0);
        if constexpr (K == 1) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi, b.lo, acc, 0, 0, 0);
        if constexpr (K == 2) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi, b.hi, acc, 0, 0, 0);
    } else {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.h, b.h, acc, 0, 0, 0);
    }
}
// One slice into two row tiles, the products interleaved so that no MFMA
// waits on the one before it.
template <int P, int K = 0>
__device__ __forceinline__ void mma_pair(const Left<P>& a0, const Left<P>& a1, const Slice<P>& b, f16v& acc0,
                                         f16v& acc1)
{
    mma_part<P, K>(a0, b, acc0);
    mma_part<P, K>(a1, b, acc1);
    if constexpr (K + 1 < nprod<P>()) mma_pair<P, K + 1>(a0, a1, b, acc0, acc1);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// e^{-2 pi i k x / N} for integer k and FP32 x: the phase k x / N is reduced
// mod 1 in FP64 (k x is exact there), then split t = th + tl so that the FP32
// sincospi of th is corrected to first order by tl.
__device__ __forceinline__ float2 expi_neg(int k, float x, double invN)
{
    double t = (double)k * (double)x * invN;
    t -= rint(t);
    const float th = (float)t;
    const float tl = (float)(t - (double)th);
    float s, c;
    sincospif(2.0f * th, &s, &c);
    const float d = 6.28318530717958647692f * tl;
    return make_float2(fmaf(-d, s, c), -fmaf(d, c, s));
}

// Right operand G, slice (ch, ks, t, lane) -- tile (ch, ks, t) of 64 lane
// slices, stored part-major (see the end of the kernel): element e of lane (h = lane / 32,
// cc = lane % 32) is G[K = 16 ks + 8 h + e][col = ch * 32 C + 32 t + cc] with
//   k2 index j2 = 8 ks + 4 h + e / 2, component e % 2 (0: Re B, 1: Im B),
//   k1 index j1 = ch * 16 C + 32 (t / 2) + cc, part t % 2 (0: Re T, 1: Im T):
//   Re T = sum Re B Re F - Im B Im F,  Im T = sum Re B Im F + Im B Re F.
// Entries with j1 >= N1 or j2 >= N2 are zero (padding to whole tiles).
template <int P, int C>
__global__ __launch_bounds__(256) void k_ndft_prep(const float2* __restrict__ F, const float* __restrict__ rowscale,
                                                   Slice<P>* __restrict__ G, int N1, int N2, int nch, int nks)
{
    const long long total = (long long)nch * nks * C * 64;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int lane = (int)(i & 63);
    long long r = i >> 6;
    const int t = (int)(r % C);
    r /= C;
    const int ks = (int)(r % nks);
    const int ch = (int)(r / nks);
    const int h = lane >> 5, cc = lane & 31;
    const int j1 = ch * 16 * C + 32 * (t >> 1) + cc;
    const int part = t & 1;
    const float sc = j1 < N1 ? rowscale[j1] : 1.0f;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j2 = 8 * ks + 4 * h + (e >> 1), comp = e & 1;
        float2 f = make_float2(0.f, 0.f);
        if (j1 < N1 && j2 < N2) f = F[(size_t)j1 * N2 + j2];
        f.x *= sc;
        f.y *= sc;
        v[e] = part == 0 ? (comp == 0 ? f.x : -f.y) : (comp == 0 ? f.y : f.x);
    }
    Slice<P> s;
    if constexpr (P == kNdftFp32) {
#pragma unroll
        for (int e = 0; e < 8; ++e) s.v[e] = v[e];
    } else if constexpr (P == kNdftBf16x3) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            __bf16 hh, ll;
            split_bf16(v[e], hh, ll);
            s.hi[e] = hh;
            s.lo[e] = ll;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) s.h[e] = (_Float16)v[e];
    }
    // a tile stores its lanes' slices as 16-byte parts, part-major:
    // word (tile * kParts + part) * 64 + lane, so that one ds_read_b128 of a
    // part by 64 lanes reads 1 KB of consecutive LDS
    constexpr int kParts = (int)sizeof(Slice<P>) / 16;
    const uint4* sw = reinterpret_cast<const uint4*>(&s);
    uint4* Gw = reinterpret_cast<uint4*>(G);
    const long long tile = i >> 6;
#pragma unroll
    for (int q = 0; q < kParts; ++q) Gw[(tile * kParts + q) * 64 + lane] = sw[q];
}

// Slice (tile, lane) of a part-major tile array in LDS.
template <int P>
__device__ __forceinline__ Slice<P> load_slice(const uint4* base, int tile, int lane)
{
    constexpr int kParts = (int)sizeof(Slice<P>) / 16;
    Slice<P> s;
    uint4* sw = reinterpret_cast<uint4*>(&s);
#pragma unroll
    for (int q = 0; q < kParts; ++q) sw[q] = base[(tile * kParts + q) * 64 + lane];
    return s;
}

// Per-row scale of F for the FP16 operand: a power of two that brings the
// row's largest |Re| or |Im| to [2^13, 2^14).  One block per row; rows with no
// finite non-zero entry get 1.  The other precisions use 1.
__global__ __launch_bounds__(256) void k_ndft_rowscale(const float2* __restrict__ F, float* __restrict__ scale,
                                                       float* __restrict__ inv_scale, int N2, int fp16)
{
    __shared__ float red[256];
    const int j1 = blockIdx.x;
    float mx = 0.f;
    if (fp16)
        for (int j2 = threadIdx.x; j2 < N2; j2 += 256) {
            const float2 f = F[(size_t)j1 * N2 + j2];
            mx = fmaxf(mx, fmaxf(fabsf(f.x), fabsf(f.y)));
        }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float m = red[0];
        int e = 0;
        if (m > 0.f && m <= 3.0e38f) {
            int ex;
            frexpf(m, &ex);           // m = f 2^ex, f in [0.5, 1)
            e = 14 - ex;              // m 2^e in [2^13, 2^14)
        }
        scale[j1] = ldexpf(1.0f, e);
        inv_scale[j1] = ldexpf(1.0f, -e);
    }
}

// The plan pads the K-step count to a multiple of this with zero slices.
constexpr int kKsPad = 2;
// Column tiles per chunk: 8 tiles = 4 groups of 32 k1 = 128 k1 per pass over K.
constexpr int kChunkTiles = 8;
constexpr int kChunkK1 = 16 * kChunkTiles;
// Row tiles (of 32 points) per wave: each right-operand slice read from LDS
// feeds both.
constexpr int kRowTiles = 2;
constexpr int kPtsPerWave = 32 * kRowTiles;
constexpr int kPtsPerGroup = kPtsPerWave * kWaves;   // points per workgroup

// Asynchronous copy of one 16-byte word per lane from global memory into LDS:
// lane i of the wave writes lds_base + 16 i (global_load_lds_dwordx4).
__device__ __forceinline__ void copy16_to_lds(const uint4* gsrc, uint4* lds_base)
{
    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)gsrc,
                                     (__attribute__((address_space(3))) void*)lds_base, 16, 0, 0);
}

__device__ __forceinline__ void wait_vm_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Main contraction.  Workgroup = 4 waves, one per SIMD; wave w owns points
// 64 w .. 64 w + 63 of the group as two MFMA row tiles of 32.  For each chunk
// of 128 k1 (8 column tiles: Re and Im tiles of 4 groups of 32 k1) the wave
// runs the K loop over all k2 with 2 x 8 accumulator tiles (256 registers per
// lane): per K-step it generates the point operands of its two row tiles and
// multiplies each right-operand slice, read once from LDS, into both.  The
// slices of the next K-step are copied by the workgroup from global memory
// straight into LDS while the current ones are multiplied (two buffers, one
// barrier per K-step).  Accumulator element r of lane (h, cc) holds row
// 8 (r / 4) + 4 h + r % 4, column cc.  After the K loop each (row tile, group
// of 32 k1) is transposed through LDS so that lane (h, row) folds columns
// 16 h .. 16 h + 15 of its row with A[m][k1] = e^{-2 pi i k1 x1_m / N1}, A
// advancing by e^{-2 pi i x1_m / N1} per column from an exact value every
// EANCHOR columns.  The point operand of every K-step is evaluated exactly
// (one phase reduction, then three products with the k2 step phasors).
template <int P, int EANCHOR>
__global__ __launch_bounds__(256, 1) void k_ndft_t2(const Slice<P>* __restrict__ G, const float* __restrict__ inv_scale,
                                                    const float* __restrict__ x1, const float* __restrict__ x2,
                                                    float2* __restrict__ out, int N1, int N2, int M, int nch, int nks)
{
    constexpr int C = kChunkTiles, R = kRowTiles;
    constexpr int kWords = C * 64 * (int)sizeof(Slice<P>) / 16;      // 16-byte words per K-step
    constexpr int kWordsPerThread = kWords / 256;
    static_assert(kWords % 256 == 0, "a K-step must split evenly over the workgroup");
    // two distinct buffers (not one indexed array), so that the compiler can
    // tell reads of one from the copy in flight into the other
    __shared__ uint4 s_stage0[kWords];
    __shared__ uint4 s_stage1[kWords];
    __shared__ float2 s_t[kWaves][32][33];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int h = lane >> 5, cc = lane & 31;
    const int m0 = blockIdx.x * kPtsPerGroup + kPtsPerWave * w;   // first point of this wave
    const double invN1 = 1.0 / N1, invN2 = 1.0 / N2;
    const int half1 = N1 / 2, half2 = N2 / 2;
    const int total = nch * nks;
    const uint4* Gw = reinterpret_cast<const uint4*>(G);

    // K-step 0 in flight while the phasors are set up
#pragma unroll
    for (int i = 0; i < kWordsPerThread; ++i)
        copy16_to_lds(Gw + 256 * i + 64 * w + lane, &s_stage0[256 * i + 64 * w]);

    // point of row cc of row tile r (left operand, and row cc in the epilogue)
    float xr1[R], xr2[R];
    float2 w1[R], w2[R], w3[R], e1[R], y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int m = m0 + 32 * r + cc;
        xr2[r] = m < M ? x2[m] : 0.f;
        xr1[r] = m < M ? x1[m] : 0.f;
        w1[r] = expi_neg(1, xr2[r], invN2);
        w2[r] = expi_neg(2, xr2[r], invN2);
        w3[r] = expi_neg(3, xr2[r], invN2);
        e1[r] = expi_neg(1, xr1[r], invN1);
        y[r] = make_float2(0.f, 0.f);
    }
    wait_vm_all();
    __syncthreads();

    f16v acc[R][C];
    // point operand of row tile r at K-step ks: e^{-2 pi i k2 x2 / N2} for
    // k2 = 8 ks + 4 h + 0..3 - N2/2, (Re, Im) interleaved
    auto gen = [&](int r, int ks) {
        float2 z[4];
        z[0] = expi_neg(8 * ks + 4 * h - half2, xr2[r], invN2);
        z[1] = cmul(z[0], w1[r]);
        z[2] = cmul(z[0], w2[r]);
        z[3] = cmul(z[0], w3[r]);
        float v[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[2 * e] = z[e].x;
            v[2 * e + 1] = z[e].y;
        }
        Left<P> a;
        make_left<P>(v, a);
        return a;
    };
    Left<P> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = gen(r, 0);
    // one K-step: multiply the slices in `rd` with a[] while the next K-step is
    // copied into `nx` and its point operands are generated
    auto step = [&](int gs, int ks, const uint4* rd, uint4* nx) {
        {
            // the last K-step copies its own slices again into the idle buffer
            // (no branch in the loop body)
            const uint4* src = Gw + (size_t)min(gs + 1, total - 1) * kWords + 64 * w + lane;
#pragma unroll
            for (int i = 0; i < kWordsPerThread; ++i) copy16_to_lds(src + 256 * i, nx + 256 * i + 64 * w);
        }
        Slice<P> b[C];
#pragma unroll
        for (int t = 0; t < C; ++t) b[t] = load_slice<P>(rd, t, lane);
#pragma unroll
        for (int t = 0; t < C; ++t) mma_pair<P>(a[0], a[1], b[t], acc[0][t], acc[1][t]);
        const int kn = ks + 1 == nks ? 0 : ks + 1;
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = gen(r, kn);
        // schedule: the reads of 2 tiles, then per tile the reads of the tile
        // two ahead, the tile's MFMAs (both row tiles) and a share of the next
        // K-step's point-operand arithmetic, so that each MFMA group waits only
        // for reads issued a group earlier and the arithmetic runs under the
        // MFMAs (sched_group_barrier masks: 0x100 LDS read, 0x008 MFMA, 0x002 VALU)
        constexpr int kR = (int)sizeof(Slice<P>) / 16, kQ = nprod<P>();
        __builtin_amdgcn_sched_group_barrier(0x100, 2 * kR, 0);
#pragma unroll
        for (int t = 0; t < C; ++t) {
            if (t + 2 < C) __builtin_amdgcn_sched_group_barrier(0x100, kR, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 2 * kQ, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 28, 0);
        }
        wait_vm_all();
        __syncthreads();
    };
    // nks is even, so K-step ks of every chunk reads buffer ks % 2
    for (int ch = 0; ch < nch; ++ch) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int t = 0; t < C; ++t) acc[r][t] = f16v{};
        for (int ks = 0; ks < nks; ks += 2) {
            step(ch * nks + ks, ks, s_stage0, s_stage1);
            step(ch * nks + ks + 1, ks + 1, s_stage1, s_stage0);
        }
        // epilogue of chunk ch: y += sum_k1 A[m][k1] T[m][k1] over its 128 k1
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int g = 0; g < C / 2; ++g) {
                const int j1c = ch * kChunkK1 + 32 * g + cc;        // k1 index of column cc
                const float isc = (P == kNdftFp16 && j1c < N1) ? inv_scale[j1c] : 1.0f;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int ti = 8 * (q >> 2) + 4 * h + (q & 3);
                    s_t[w][ti][cc] = make_float2(acc[r][2 * g][q] * isc, acc[r][2 * g + 1][q] * isc);
                }
                wave_sync();
                const int k1s = ch * kChunkK1 + 32 * g + 16 * h - half1;   // k1 of this lane's first column
                float2 A = make_float2(1.f, 0.f);
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    if (c % EANCHOR == 0) A = expi_neg(k1s + c, xr1[r], invN1);
                    const float2 tv = s_t[w][cc][16 * h + c];
                    y[r].x = fmaf(A.x, tv.x, fmaf(-A.y, tv.y, y[r].x));
                    y[r].y = fmaf(A.x, tv.y, fmaf(A.y, tv.x, y[r].y));
                    if (c % EANCHOR != EANCHOR - 1) A = cmul(A, e1[r]);
                }
                wave_sync();
            }
        }
    }
    // the two halves of each row
#pragma unroll
    for (int r = 0; r < R; ++r) {
        float2 v = y[r];
        v.x += __shfl_xor(v.x, 32, 64);
        v.y += __shfl_xor(v.y, 32, 64);
        const int m = m0 + 32 * r + cc;
        if (h == 0 && m < M) out[m] = v;
    }
}

}  // namespace ndft
