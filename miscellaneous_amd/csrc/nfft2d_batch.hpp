// nfft2d_batch.hpp -- many type-2 NFFTs of different sizes in one launch per stage.
//
// A batch holds one plan per transform (complex64, K = 6, c = 2, N1 and N2
// even, up to 1024; powers of two from 8 run the Stockham passes, other sizes
// the direct ones).  Every stage of setpts and of the type-2 runs as
// ONE kernel over the blocks of all transforms: block b finds its transform
// and its block index inside the transform in an owner table built when the
// batch is created (the block counts of these stages depend on the sizes
// only).  The gather is persistent: its blocks walk one list of the
// subproblems of all transforms, each entry tagged with its transform.
// Together: 5 launches for setpts, 4 for the type-2 and 4 for the type-1,
// whatever the number of transforms, all stream-ordered and free of
// allocation, so the whole sequence can be captured into one hipGraph.
#pragma once

#include "nfft2d_czt.hpp"
#include "nfft2d_kernels.hpp"

namespace nfft2d {

// Per-transform state, device-resident (one array per batch).
struct BatchDesc {
    int N1, N2, L1, L2, nb1, nbins, M, maxpts, nblk, chunk, ngroups;
    float inv1, inv2;
    BinGeom geo;
    // plan buffers
    float2 *Tm, *Gt;
    const float *D1, *D2;
    const float2 *tw1, *tw2;
    int *hist, *rank, *tot, *subbeg, *nsub;
    int2 *gexcl, *gsum;
    int4* sub;
    PtRec<float>* pts;
    // the caller's arrays
    const float *x, *y;
    const float2* f;
    float2* out;
    // type-1: samples in, modes out; spread tiles; tile-merge radii (merge_radius)
    const float2* d1;
    float2* F1;
    float2* scratch;
    int r1, r2;
    // type-2 gather: subproblems of at most dmax points are read point by point
    // from the grid (kb_gather_direct), larger ones through an LDS tile
    int dmax;
    // chirp-z tables of the plan per dimension (null: Stockham or direct passes)
    const void *czt1, *czt2;
    int cztm1, cztm2;
};

// Largest N of the batched FFT passes.
constexpr int kBatchMaxN = 1024;

// Powers of two from 8 run the Stockham passes; other even N the direct ones.
__host__ __device__ constexpr bool batch_fft_n(int N) { return N >= 8 && (N & (N - 1)) == 0; }

template <typename F>
__device__ __forceinline__ void batch_dispatch_n(int N, F&& f)
{
    switch (N) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    case 1024: f(std::integral_constant<int, 1024>{}); break;
    default: break;
    }
}

template <int N> constexpr int cmax_rows_lds() { return t2_rows_lds<float, N>(); }
constexpr int batch_rows_lds()
{
    int m = 0;
    m = m > cmax_rows_lds<8>() ? m : cmax_rows_lds<8>();
    m = m > cmax_rows_lds<16>() ? m : cmax_rows_lds<16>();
    m = m > cmax_rows_lds<32>() ? m : cmax_rows_lds<32>();
    m = m > cmax_rows_lds<64>() ? m : cmax_rows_lds<64>();
    m = m > cmax_rows_lds<128>() ? m : cmax_rows_lds<128>();
    m = m > cmax_rows_lds<256>() ? m : cmax_rows_lds<256>();
    m = m > cmax_rows_lds<512>() ? m : cmax_rows_lds<512>();
    m = m > cmax_rows_lds<1024>() ? m : cmax_rows_lds<1024>();
    return m;
}
template <int N> constexpr int cmax_cols_lds() { return t2_cols_lds<float, N>(); }
constexpr int batch_cols_lds()
{
    int m = 0;
    m = m > cmax_cols_lds<8>() ? m : cmax_cols_lds<8>();
    m = m > cmax_cols_lds<16>() ? m : cmax_cols_lds<16>();
    m = m > cmax_cols_lds<32>() ? m : cmax_cols_lds<32>();
    m = m > cmax_cols_lds<64>() ? m : cmax_cols_lds<64>();
    m = m > cmax_cols_lds<128>() ? m : cmax_cols_lds<128>();
    m = m > cmax_cols_lds<256>() ? m : cmax_cols_lds<256>();
    m = m > cmax_cols_lds<512>() ? m : cmax_cols_lds<512>();
    m = m > cmax_cols_lds<1024>() ? m : cmax_cols_lds<1024>();
    return m;
}

// ---------------------------------------------------------------- setpts
__global__ __launch_bounds__(kHistThreads) void kb_hist(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    extern __shared__ int h[];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    // no per-point rank: kb_scatter_lds ranks the points again in LDS
    hist_body<float>(o.y, h, d.x, d.y, d.M, d.chunk, d.L1, d.L2, d.inv1, d.inv2, d.nb1, d.nbins, d.hist, nullptr);
}

__global__ __launch_bounds__(256) void kb_colscan(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    colscan_body(o.y, d.hist, d.nblk, d.nbins, d.maxpts, d.tot, d.gexcl, d.gsum);
}

// one block per transform
__global__ __launch_bounds__(kScanThreads) void kb_scan_groups(const BatchDesc* __restrict__ bd)
{
    const BatchDesc& d = bd[blockIdx.x];
    scan_groups_body(d.gsum, d.ngroups, d.nbins, d.subbeg, d.nsub);
}

// (one block) base[t] = sum of nsub over the transforms before t; base[count] = total
__global__ __launch_bounds__(1024) void kb_sub_prefix(const BatchDesc* __restrict__ bd, int count, int* __restrict__ base)
{
    __shared__ int w_sum[16], carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < count; b0 += 1024) {
        const int i = b0 + t;
        const int v = i < count ? *bd[i].nsub : 0;
        int a = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int n = __shfl_up(a, o, 64);
            if (lane >= o) a += n;
        }
        if (lane == 63) w_sum[w] = a;
        __syncthreads();
        int pre = carry + a - v;
        for (int k = 0; k < w; ++k) pre += w_sum[k];
        if (i < count) base[i] = pre;
        __syncthreads();
        if (t == 1023) carry = pre + v;
        __syncthreads();
    }
    if (t == 0) base[count] = carry;
}

// Scatter of every transform's points, and each transform's subproblems into
// the batch list: entry base[t] + s = (bin, start, count, t), with ~t (< 0) in
// place of t when count <= the transform's direct-gather cut-off dmax.
// Block (t, k) takes the points of histogram block k of transform t (the
// histogram stage's owner table) and, as k_scatter_lds does for a plan,
// builds pos[bin] = first sorted position of (block k, bin) in LDS from its row
// of column offsets, then gives each point pos[bin] + its rank by incrementing
// pos[bin] (one LDS atomic per run of equal bins in a wave).  So kb_hist
// stores no per-point rank (4 bytes per point written and read back), and no
// point's position waits on a scattered global read of the offsets.  The
// transform's bins' subproblem lists are split evenly over its blocks.
__global__ __launch_bounds__(kScatThreads) void kb_scatter_lds(const BatchDesc* __restrict__ bd,
                                                               const int2* __restrict__ own,
                                                               const int* __restrict__ base, int4* __restrict__ bsub)
{
    extern __shared__ int pos_lds[];                         // the transform's nbins (at most kMaxLdsBins)
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    const int t = threadIdx.x;
    const int lo = o.y * d.chunk, hi = min(lo + d.chunk, d.M);
    // the first batch of points: loads in flight while the positions are built
    float xv[kScatPer], yv[kScatPer];
#pragma unroll
    for (int q = 0; q < kScatPer; ++q) {
        const int i = lo + q * kScatThreads + t;
        xv[q] = i < hi ? d.x[i] : 0.f;
        yv[q] = i < hi ? d.y[i] : 0.f;
    }
    const int* __restrict__ orow = d.hist + (size_t)o.y * d.nbins;
    for (int b = t; b < d.nbins; b += kScatThreads) pos_lds[b] = d.gsum[b / kColBins].x + d.gexcl[b].x + orow[b];
    const int bpb = (d.nbins + d.nblk - 1) / d.nblk;
    const int blo = o.y * bpb, bhi = min(blo + bpb, d.nbins);
    for (int i = blo + t; i < bhi; i += kScatThreads) {
        const int c = d.tot[i];
        const int2 e = d.gexcl[i], gb = d.gsum[i / kColBins];
        const int pos = gb.x + e.x, sp = base[o.x] + gb.y + e.y;
        d.subbeg[i] = gb.y + e.y;                            // the bin's first subproblem in its transform
        const int ns = (c + d.maxpts - 1) / d.maxpts;
        for (int k = 0; k < ns; ++k) {
            const int n = min(d.maxpts, c - k * d.maxpts);
            // the transform index, bit-complemented for the subproblems the type-2 gathers lane by lane
            bsub[sp + k] = make_int4(i, pos + k * d.maxpts, n, n <= d.dmax ? ~o.x : o.x);
        }
    }
    __syncthreads();
    for (int b0 = lo; b0 < hi; b0 += kScatPer * kScatThreads) {
        if (b0 != lo) {
#pragma unroll
            for (int q = 0; q < kScatPer; ++q) {
                const int i = b0 + q * kScatThreads + t;
                xv[q] = i < hi ? d.x[i] : 0.f;
                yv[q] = i < hi ? d.y[i] : 0.f;
            }
        }
#pragma unroll
        for (int q = 0; q < kScatPer; ++q) {
            const int i = b0 + q * kScatThreads + t;
            const bool act = i < hi;
            int m1, m2;
            float t1, t2;
            const int bin = bin_of(xv[q], yv[q], d.L1, d.L2, d.inv1, d.inv2, d.nb1, m1, m2, t1, t2);
            const int p = wave_run_rank(pos_lds, bin, act);
            if (act) {
                PtRec<float> rec;
                rec.t1 = t1;
                rec.t2 = t2;
                rec.cell = (m2 << 16) | m1;
                rec.idx = i;
                d.pts[p] = rec;
            }
        }
    }
}

// ---------------------------------------------------------------- grid passes
// Row passes: 128 threads, block o.y of the transform's rows (R rows of N2 for
// the Stockham passes, one row for the chirp-z and direct ones); column
// passes: 256 threads, likewise over the L2 columns.  The LDS buffer fits the
// largest N of every form: the direct passes stage one line of up to
// 2 kBatchMaxN, the chirp-z passes (the plans' for lengths up to
// kBatchCztMaxN; longer ones take the direct passes here) one FFT of up to
// 2048 points with its twiddles.
constexpr int kBatchCztMaxM = 2048;
constexpr int kBatchCztU = 4;   // values per thread: 4 U >= kBatchCztMaxM / 128
constexpr int kBatchLineLds = 2 * kBatchMaxN > czt_lds_entries(kBatchCztMaxM) ? 2 * kBatchMaxN
                                                                             : czt_lds_entries(kBatchCztMaxM);
constexpr int batch_lds_rows() { return batch_rows_lds() > kBatchLineLds ? batch_rows_lds() : kBatchLineLds; }
constexpr int batch_lds_cols() { return batch_cols_lds() > kBatchLineLds ? batch_cols_lds() : kBatchLineLds; }

__global__ __launch_bounds__(128) void kb_t2_rows(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[batch_lds_rows()];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    if (!batch_fft_n(d.N2)) {
        if (d.czt2)   // row o.y: f[k1][*] (N2) -> Tm[k1][*] (L2), pre-compensated
            czt_line<float, -1, true, kBatchCztU>(o.y, lds, d.f, d.N2, 1, d.Tm, 2 * d.N2, 1, d.D1, d.D2,
                                                  czt_tables_of<float>(d.czt2, d.N2, d.cztm2, true));
        else
            t2_rows_direct_body<float, -1>(o.y, lds, d.f, d.Tm, d.D1, d.D2, d.tw2, d.N1, d.N2,
                                           3 * d.N2 <= batch_lds_rows() ? lds + d.N2 : nullptr);
        return;
    }
    batch_dispatch_n(d.N2, [&](auto n) {
        constexpr int N = decltype(n)::value;
        t2_rows_body<float, N, -1>(o.y, lds, d.f, d.Tm, d.D1, d.D2, d.tw2, d.N1);
    });
}

__global__ __launch_bounds__(256) void kb_t2_cols(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[batch_lds_cols()];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    if (!batch_fft_n(d.N1)) {
        if (d.czt1)   // column o.y: Tm[*][n2] (N1) -> Gt[n2][*] (L1)
            czt_line<float, -1, true, kBatchCztU>(o.y, lds, d.Tm, 1, d.L2, d.Gt, d.L1, 1, (const float*)nullptr,
                                                  (const float*)nullptr,
                                                  czt_tables_of<float>(d.czt1, d.N1, d.cztm1, true));
        else
            t2_cols_direct_body<float, -1>(o.y, lds, d.Tm, d.Gt, d.tw1, d.N1, d.L2,
                                           3 * d.N1 <= batch_lds_cols() ? lds + d.N1 : nullptr);
        return;
    }
    batch_dispatch_n(d.N1, [&](auto n) {
        constexpr int N = decltype(n)::value;
        t2_cols_body<float, N, -1>(o.y, lds, d.Tm, d.Gt, d.tw1, d.L2);
    });
}

// ---------------------------------------------------------------- type-1
// Spread: one block per entry of the batch's subproblem list (base[count]
// entries), the plan kernel's split-FP16 MFMA tiles into each transform's
// own tile scratch, indexed by the subproblem's place in its transform.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void kb_spread(
    const BatchDesc* __restrict__ bd, const int4* __restrict__ bsub, const int* __restrict__ base, int count, WinArgs wa)
{
    const int s = blockIdx.x;
    if (s >= base[count]) return;
    const int4 e = bsub[s];
    const int tr = e.w < 0 ? ~e.w : e.w;
    const BatchDesc& t = bd[tr];
    spread_halo_body<false>(s - base[tr], make_int4(e.x, e.y, e.z, 0), t.pts, t.d1, t.subbeg, t.tot, t.maxpts, t.geo,
                            wa, t.scratch);
}

// Merge: one block per bin of each transform (owner table), summing the
// neighbouring tiles as k_spread_merge does.
__global__ __launch_bounds__(256) void kb_spread_merge(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    const int2 o = own[blockIdx.x];
    const BatchDesc& t = bd[o.x];
    const int nb2 = t.nbins / t.geo.nb1;
    if (t.r1 == 1 && t.r2 == 1) spread_merge_body<float, 1, 1>(o.y, t.scratch, t.subbeg, t.geo, nb2, t.Gt);
    else if (t.r1 == 1) spread_merge_body<float, 1, 2>(o.y, t.scratch, t.subbeg, t.geo, nb2, t.Gt);
    else if (t.r2 == 1) spread_merge_body<float, 2, 1>(o.y, t.scratch, t.subbeg, t.geo, nb2, t.Gt);
    else spread_merge_body<float, 2, 2>(o.y, t.scratch, t.subbeg, t.geo, nb2, t.Gt);
}

__global__ __launch_bounds__(256) void kb_t1_cols(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[batch_lds_cols()];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    if (!batch_fft_n(d.N1)) {
        if (d.czt1)   // column o.y: Gt[n2][*] (L1) -> Tm[*][n2] (N1)
            czt_line<float, -1, false, kBatchCztU>(o.y, lds, d.Gt, d.L1, 1, d.Tm, 1, d.L2, (const float*)nullptr,
                                                   (const float*)nullptr,
                                                   czt_tables_of<float>(d.czt1, d.N1, d.cztm1, false));
        else
            t1_cols_direct_body<float, -1>(o.y, lds, d.Gt, d.Tm, d.tw1, d.N1, d.L2,
                                           4 * d.N1 <= batch_lds_cols() ? lds + 2 * d.N1 : nullptr);
        return;
    }
    batch_dispatch_n(d.N1, [&](auto n) {
        constexpr int N = decltype(n)::value;
        t1_cols_body<float, N, -1>(o.y, lds, d.Gt, d.Tm, d.tw1, d.L2);
    });
}

__global__ __launch_bounds__(128) void kb_t1_rows(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[batch_lds_rows()];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    if (!batch_fft_n(d.N2)) {
        if (d.czt2)   // row o.y: Tm[k1][*] (L2) -> F1[k1][*] (N2), deconvolved
            czt_line<float, -1, false, kBatchCztU>(o.y, lds, d.Tm, 2 * d.N2, 1, d.F1, d.N2, 1, d.D1, d.D2,
                                                   czt_tables_of<float>(d.czt2, d.N2, d.cztm2, false));
        else
            t1_rows_direct_body<float, -1>(o.y, lds, d.Tm, d.F1, d.D1, d.D2, d.tw2, d.N1, d.N2,
                                           4 * d.N2 <= batch_lds_rows() ? lds + 2 * d.N2 : nullptr);
        return;
    }
    batch_dispatch_n(d.N2, [&](auto n) {
        constexpr int N = decltype(n)::value;
        t1_rows_body<float, N, -1>(o.y, lds, d.Tm, d.F1, d.D1, d.D2, d.tw2, d.N1);
    });
}

// ---------------------------------------------------------------- type-1 direct passes on MFMA
// The matrix-core passes (t1_cols_mm_compute, t1_rows_mm_compute) over a
// batch: one block per 16-line tile of a transform, listed in the owner
// tables.  The direct passes on the VALU read a line value and a twiddle from
// LDS for every 4 FMAs and ran at ~0.19 of the VALU peak.  Sums run in another
// order than the VALU direct passes: a type-1 in a batch is held to its
// plan's result within rounding, not to its bits.
__host__ __device__ inline bool t1_pass_on_mfma(int N) { return !batch_fft_n(N) && N <= kT1MmMaxN; }
// The matrix-core column pass sums the spread tiles itself (and the tile merge
// skips the transform) when the merge needs the 3 x 3 neighbours only.
__host__ __device__ inline bool t1_cols_merge_tiles(int N1, int r1, int r2)
{
    return t1_pass_on_mfma(N1) && r1 == 1 && r2 == 1;
}

// Column pass, column tile o.y of transform o.x: Tm[i1][n2] for the 16
// columns n2 of the tile, every row tile of i1.  When the merge needs the
// 3 x 3 neighbours only, the 16 lines (one bin row of the grid) are summed
// here from the spread tiles that cover each cell (merge_cell, the tile
// merge's sum in its order), so these transforms skip kb_spread_merge and the
// grid's write and re-read.
template <int R1, int R2>
__device__ __forceinline__ void mm_stage_merged(const BatchDesc& d, int ct, float2* panel, int L1, int L2)
{
    const int nb2 = d.nbins / d.geo.nb1;
    for (int e = threadIdx.x; e < 16 * L1; e += blockDim.x) {
        const int j = e / L1, n1 = e - j * L1, n2 = 16 * ct + j;
        panel[j * kT1MmLd + n1] =
            n2 < L2 ? merge_cell<float, R1, R2>(n2, n1, d.scratch, d.subbeg, d.geo, nb2) : make_float2(0.f, 0.f);
    }
}
__global__ __launch_bounds__(256) void kb_t1_cols_mm(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[kT1MmLds];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    const int N1 = d.N1, L1 = 2 * N1, L2 = d.L2;
    float2* panel = lds;                                     // [16][kT1MmLd]: rows n2 of the grid
    float2* wt = lds + 16 * kT1MmLd;                         // L1 twiddles
    if (t1_cols_merge_tiles(d.N1, d.r1, d.r2)) {
        mm_stage_merged<1, 1>(d, o.y, panel, L1, L2);
        for (int k = threadIdx.x; k < L1; k += blockDim.x) wt[k] = d.tw1[k];
    } else {   // grids with a thin last bin (5 x 5 neighbours): the merged grid from kb_spread_merge
        mm_stage_rows(d.Gt, 16 * o.y, L2, L1, d.tw1, panel, wt);
    }
    __syncthreads();
    t1_cols_mm_compute<-1>(o.y, panel, wt, d.Tm, N1, L2);
}

// Row pass, row tile o.y of transform o.x: F1[i1][i2] for the 16 rows i1 of
// the tile, every column tile of i2, deconvolved.
__global__ __launch_bounds__(256) void kb_t1_rows_mm(const BatchDesc* __restrict__ bd, const int2* __restrict__ own)
{
    __shared__ float2 lds[kT1MmLds];
    const int2 o = own[blockIdx.x];
    const BatchDesc& d = bd[o.x];
    mm_stage_rows(d.Tm, 16 * o.y, d.N1, 2 * d.N2, d.tw2, lds, lds + 16 * kT1MmLd);
    __syncthreads();
    t1_rows_mm_compute<-1>(o.y, lds, lds + 16 * kT1MmLd, d.F1, d.D1, d.D2, d.N1, d.N2);
}

#ifndef NFFT2D_BATCH_GATHER_WAVES
#define NFFT2D_BATCH_GATHER_WAVES 4
#endif
// Subproblems of at most BatchDesc::dmax points (per transform, at most
// kDirectCap) are gathered by kb_gather_direct, kDirectLanes lanes per list
// entry, the others by kb_gather_w.  A fixed cut-off of 8 gave, on the batch
// of 256 uneven transforms, 1.89 ms per setpts + type-2 against 1.91 with 4,
// 1.97 with 16 and 2.03 with every subproblem on the waves; the batch's
// autotuner picks the cut-off per size class.
constexpr int kDirectLanes = 8;
constexpr int kDirectCap = 16;
constexpr int kDirectDefault = 8;

// Persistent gather over the batch's subproblem list (total = base[count]),
// one wave per subproblem: wave g = 4 blockIdx.x + w takes subproblems g,
// g + 4 gridDim.x, ... (those of more than their transform's dmax points,
// marked by a non-negative transform index in the list).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NFFT2D_BATCH_GATHER_WAVES, 8))) void
kb_gather_w(const BatchDesc* __restrict__ bd, const int4* __restrict__ bsub, const int* __restrict__ total, WinArgs wa)
{
    __shared__ WaveTileLds lds[4];
    const int w = threadIdx.x >> 6;
    const int n = *total;
    for (int s = blockIdx.x * 4 + w; s < n; s += gridDim.x * 4) {
        const int4 d = bsub[s];
        if (d.w < 0) continue;                           // kb_gather_direct's
        const BatchDesc& t = bd[d.w];
        gather_wave_single(d, lds[w], t.Gt, t.geo, t.pts, wa, t.out);
    }
}

// The subproblems of at most dmax points: kDirectLanes threads per list entry,
// thread q computing points q and q + kDirectLanes of the entry with
// direct_point (the same bits as the wave's sums).
__global__ __launch_bounds__(256) void kb_gather_direct(const BatchDesc* __restrict__ bd, const int4* __restrict__ bsub,
                                                        const int* __restrict__ total, WinArgs wa)
{
    const int s = blockIdx.x * (256 / kDirectLanes) + (int)threadIdx.x / kDirectLanes;
    const int q = (int)threadIdx.x % kDirectLanes;
    if (s >= *total) return;
    const int4 d = bsub[s];
    if (d.w >= 0 || q >= d.z) return;                    // kb_gather_w's, or past the entry's points
    const BatchDesc& t = bd[~d.w];
    const int O1 = (d.x % t.geo.nb1) * kBin, O2 = (d.x / t.geo.nb1) * kBin;
    for (int k = q; k < d.z; k += kDirectLanes) {
        const PtRec<float> rec = t.pts[d.y + k];
        t.out[rec.idx] = direct_point(rec, O1, O2, t.Gt, t.geo, wa);
    }
}

}  // namespace nfft2d
