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
#include <type_traits>

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
// All products of one K-step: C slices into two row tiles, product part K
// outermost, so that two MFMAs into the same accumulator are 2 C - 1 MFMAs
// apart (an MFMA's result is ready about two MFMA issue slots after it starts).
template <int P, int C, int K = 0>
__device__ __forceinline__ void mma_step(const Left<P>& a0, const Left<P>& a1, const Slice<P> (&b)[C],
                                         f16v (&acc0)[C], f16v (&acc1)[C])
{
#pragma unroll
    for (int t = 0; t < C; ++t) {
        mma_part<P, K>(a0, b[t], acc0[t]);
        mma_part<P, K>(a1, b[t], acc1[t]);
    }
    if constexpr (K + 1 < nprod<P>()) mma_step<P, C, K + 1>(a0, a1, b, acc0, acc1);
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
constexpr int kKsPad = 4;
// The point operand is evaluated exactly (phase reduction + sincospi) every
// kKsExact K-steps and advanced by one complex product in between.
constexpr int kKsExact = 4;
static_assert(kKsPad % kKsExact == 0, "chunks start on an exact K-step");
// K-steps of mode slices in flight ahead of the one being multiplied: the
// L2-to-LDS copy of a K-step takes longer than the K-step's MFMAs, so the
// copies run kAhead steps ahead (kAhead + 1 LDS buffers, rotating).
constexpr int kAhead = 3;
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
// Wait until at most N vector-memory operations of this wave are in flight.
template <int N> __device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory"); }
// Workgroup barrier without a memory fence: __syncthreads' release fence
// waits for every copy in flight (vmcnt(0)), which would undo the copies that
// run ahead.  Each wave waits for its own copies with wait_vm first; the
// "memory" clobber keeps the compiler from moving LDS accesses across.
__device__ __forceinline__ void barrier_no_fence() { asm volatile("s_barrier" ::: "memory"); }

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
// EANCHOR columns.  The point operand is evaluated exactly (one phase
// reduction, then three products with the k2 step phasors) on every
// kKsExact-th K-step and advanced by e^{-2 pi i 8 x2 / N2} on the others.
//
// SPLIT = 2 runs two waves per SIMD: waves w and w + 4 hold the same points
// and split each chunk's column tiles between them (128 accumulator registers
// each instead of 256), so that one wave's VALU work and LDS waits run under
// the other wave's MFMAs; both generate the same point operands, and the two
// partial sums of a point meet in LDS at the end.
template <int P, int EANCHOR, int SPLIT>
__global__ __launch_bounds__(256 * SPLIT, 1) void k_ndft_t2(const Slice<P>* __restrict__ G,
                                                            const float* __restrict__ inv_scale,
                                                            const float* __restrict__ x1, const float* __restrict__ x2,
                                                            float2* __restrict__ out, int N1, int N2, int M, int nch,
                                                            int nks)
{
    constexpr int C = kChunkTiles, R = kRowTiles, CW = C / SPLIT, NT = 256 * SPLIT;
    constexpr int kWords = C * 64 * (int)sizeof(Slice<P>) / 16;      // 16-byte words per K-step
    constexpr int kWordsPerThread = kWords / NT;
    static_assert(kWords % NT == 0, "a K-step must split evenly over the workgroup");
    // distinct buffers (not one indexed array), so that the compiler can tell
    // reads of one from the copies in flight into the others
    static_assert(kAhead == 3, "four rotating buffers");
    __shared__ uint4 s_stage0[kWords];
    __shared__ uint4 s_stage1[kWords];
    __shared__ uint4 s_stage2[kWords];
    __shared__ uint4 s_stage3[kWords];
    __shared__ float2 s_t[kWaves * SPLIT][32][33];
    __shared__ float2 s_y[SPLIT > 1 ? kWaves * R * 32 : 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int pg = w % kWaves, hc = w / kWaves;                    // point group, column share
    const int h = lane >> 5, cc = lane & 31;
    const int m0 = blockIdx.x * kPtsPerGroup + kPtsPerWave * pg;  // first point of this wave
    const double invN1 = 1.0 / N1, invN2 = 1.0 / N2;
    const int half1 = N1 / 2, half2 = N2 / 2;
    const int total = nch * nks;
    const uint4* Gw = reinterpret_cast<const uint4*>(G);

    // copy K-step gs (clamped to the last one) into buffer dst
    auto issue = [&](int gs, uint4* dst) {
        const uint4* src = Gw + (size_t)min(gs, total - 1) * kWords + 64 * w + lane;
#pragma unroll
        for (int i = 0; i < kWordsPerThread; ++i) copy16_to_lds(src + NT * i, dst + NT * i + 64 * w);
    };
    // K-steps 0 .. kAhead - 1 in flight while the phasors are set up
    issue(0, s_stage0);
    issue(1, s_stage1);
    issue(2, s_stage2);

    // point of row cc of row tile r (left operand, and row cc in the epilogue)
    float xr1[R], xr2[R];
    float2 w1[R], w2[R], w3[R], w8[R], z0[R], e1[R], y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int m = m0 + 32 * r + cc;
        xr2[r] = m < M ? x2[m] : 0.f;
        xr1[r] = m < M ? x1[m] : 0.f;
        w1[r] = expi_neg(1, xr2[r], invN2);
        w2[r] = expi_neg(2, xr2[r], invN2);
        w3[r] = expi_neg(3, xr2[r], invN2);
        w8[r] = expi_neg(8, xr2[r], invN2);
        e1[r] = expi_neg(1, xr1[r], invN1);
        y[r] = make_float2(0.f, 0.f);
    }
    wait_vm<(kAhead - 1) * kWordsPerThread>();    // K-step 0 has landed
    barrier_no_fence();

    f16v acc[R][CW];
    // point operand of row tile r at K-step ks: e^{-2 pi i k2 x2 / N2} for
    // k2 = 8 ks + 4 h + 0..3 - N2/2, (Re, Im) interleaved
    // exact: z0 = e^{-2 pi i k2 x2 / N2} at the K-step's first k2 of this lane;
    // otherwise z0 advances by 8 k2 from the previous K-step
    auto gen = [&](int r, int ks, bool exact) {
        z0[r] = exact ? expi_neg(8 * ks + 4 * h - half2, xr2[r], invN2) : cmul(z0[r], w8[r]);
        float2 z[4];
        z[0] = z0[r];
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
    for (int r = 0; r < R; ++r) a[r] = gen(r, 0, true);
    // one K-step: multiply the slices in `rd` with a[] while K-step gs + kAhead
    // is copied into `nx` (the buffer read one step earlier; past the end the
    // last K-step is copied again, so the loop body has no branch) and the
    // next point operands are generated
    auto step = [&](int gs, int ks, const uint4* rd, uint4* nx, auto exact_next) {
        constexpr bool kExact = decltype(exact_next)::value;
        issue(gs + kAhead, nx);
        Slice<P> b[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) b[t] = load_slice<P>(rd, CW * hc + t, lane);
        mma_step<P, CW>(a[0], a[1], b, acc[0], acc[1]);
        const int kn = ks + 1 == nks ? 0 : ks + 1;
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = gen(r, kn, kExact);
        // schedule: the reads of 2 tiles; then, for product part 0, per tile
        // the reads of the tile two ahead and the tile's two MFMAs; then the
        // MFMAs of the other parts; a share of the next K-step's point-operand
        // arithmetic after each MFMA pair, so that it runs under the MFMAs
        // (sched_group_barrier masks: 0x100 LDS read, 0x008 MFMA, 0x002 VALU)
        constexpr int kR = (int)sizeof(Slice<P>) / 16, kQ = nprod<P>();
        constexpr int kV = (kExact ? 224 : 64) / (CW * kQ) + 1;
        __builtin_amdgcn_sched_group_barrier(0x100, 2 * kR, 0);
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
#pragma unroll
            for (int t = 0; t < CW; ++t) {
                if (q == 0 && t + 2 < CW) __builtin_amdgcn_sched_group_barrier(0x100, kR, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, kV, 0);
            }
        }
        // every MFMA of the step issues before the wait for the copy in flight
        // (left alone, the scheduler hoists the wait and the barrier above most
        // of the MFMAs, which then wait for the copy issued at the step's start)
        __builtin_amdgcn_sched_barrier(0);
        wait_vm<(kAhead - 1) * kWordsPerThread>();   // K-step gs + 1 has landed
        barrier_no_fence();
    };
    // nks is a multiple of 4, so K-step ks of every chunk reads buffer ks % 4
    // and the operands of K-steps 0, 4, 8, ... are the exact ones
    using Exact = std::true_type;
    using Advance = std::false_type;
    static_assert(kKsExact == 4, "the loop body below covers 4 K-steps");
    for (int ch = 0; ch < nch; ++ch) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int t = 0; t < CW; ++t) acc[r][t] = f16v{};
        for (int ks = 0; ks < nks; ks += 4) {
            const int g0 = ch * nks + ks;
            step(g0, ks, s_stage0, s_stage3, Advance{});
            step(g0 + 1, ks + 1, s_stage1, s_stage0, Advance{});
            step(g0 + 2, ks + 2, s_stage2, s_stage1, Advance{});
            step(g0 + 3, ks + 3, s_stage3, s_stage2, Exact{});
        }
        // epilogue of chunk ch: y += sum_k1 A[m][k1] T[m][k1] over the wave's
        // groups of 32 k1 of the chunk (group gg = its Re and Im tiles)
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int g = 0; g < CW / 2; ++g) {
                const int gg = (CW / 2) * hc + g;
                const int j1c = ch * kChunkK1 + 32 * gg + cc;       // k1 index of column cc
                const float isc = (P == kNdftFp16 && j1c < N1) ? inv_scale[j1c] : 1.0f;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int ti = 8 * (q >> 2) + 4 * h + (q & 3);
                    s_t[w][ti][cc] = make_float2(acc[r][2 * g][q] * isc, acc[r][2 * g + 1][q] * isc);
                }
                wave_sync();
                const int k1s = ch * kChunkK1 + 32 * gg + 16 * h - half1;  // k1 of this lane's first column
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
    // the two halves of each row, then (SPLIT = 2) the two column shares
#pragma unroll
    for (int r = 0; r < R; ++r) {
        y[r].x += __shfl_xor(y[r].x, 32, 64);
        y[r].y += __shfl_xor(y[r].y, 32, 64);
        if (SPLIT > 1 && hc == 1 && h == 0) s_y[(pg * R + r) * 32 + cc] = y[r];
    }
    if constexpr (SPLIT > 1) __syncthreads();
    if (hc == 0 && h == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 v = y[r];
            if constexpr (SPLIT > 1) {
                const float2 o = s_y[(pg * R + r) * 32 + cc];
                v.x += o.x;
                v.y += o.y;
            }
            const int m = m0 + 32 * r + cc;
            if (m < M) out[m] = v;
        }
    }
}

}  // namespace ndft
