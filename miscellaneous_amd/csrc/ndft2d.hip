// ndft2d.hip -- the 2D NDFT.  Part 1: NDFT2_2D_GPU, the reference's direct
// sum in FP64 (NFFT/NDFT2_2D.cu:15-32,37-72):
//   out[n] = sum_{m<M} in[m] exp(-2 pi i (u[n] X[m] / Nu + v[n] Y[m] / Nv)).
// The reference fills an explicit M x N FP64 kernel matrix (:44-50; 64 GiB at
// M = N = 65536), writes it to a text file (:56) and calls cublasZgemv (:67).
// Here the kernel is evaluated on the fly: each thread owns one output n and
// streams the points through LDS in chunks, accumulating in FP64.
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/nfft2d.h"
#include "ndft2d_mfma.hpp"
#include "nfft2d_internal.hpp"

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_ndft_direct(const double *__restrict__ X, const double *__restrict__ Y,
                                                          const double *__restrict__ u, const double *__restrict__ v,
                                                          const double2 *__restrict__ in, double2 *__restrict__ out,
                                                          double inv_Nu, double inv_Nv, int M, int N)
{
    __shared__ double sX[kThreads], sY[kThreads];
    __shared__ double2 sIn[kThreads];
    const int n = blockIdx.x * kThreads + threadIdx.x;
    const double un = n < N ? u[n] * inv_Nu : 0.0;
    const double vn = n < N ? v[n] * inv_Nv : 0.0;
    double ar = 0.0, ai = 0.0;
    for (int m0 = 0; m0 < M; m0 += kThreads) {
        const int m = m0 + threadIdx.x;
        if (m < M) {
            sX[threadIdx.x] = X[m];
            sY[threadIdx.x] = Y[m];
            sIn[threadIdx.x] = in[m];
        }
        __syncthreads();
        const int cnt = min(kThreads, M - m0);
        for (int q = 0; q < cnt; ++q) {
            double t = un * sX[q] + vn * sY[q];
            t -= floor(t);
            double s, c;
            sincospi(2.0 * t, &s, &c);   // e^{-2 pi i t} = c - i s
            const double2 d = sIn[q];
            ar += d.x * c + d.y * s;
            ai += d.y * c - d.x * s;
        }
        __syncthreads();
    }
    if (n < N) out[n] = make_double2(ar, ai);
}

}  // namespace

extern "C" int NDFT2_2D_GPU(void *handle, const double *X, const double *Y, const double *u, const double *v,
                            const nfft2d_c128 *in, nfft2d_c128 *out, int Nu, int Nv, int M, int N)
{
    (void)handle;
    if (M < 0 || N < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: negative size");
    if (Nu == 0 || Nv == 0) return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: Nu and Nv must be non-zero");
    if (N == 0) return NFFT2D_OK;
    if (!u || !v || !out || (M > 0 && (!X || !Y || !in)))
        return nfft2d_fail(NFFT2D_ERR_ARG, "NDFT2_2D_GPU: null pointer");
    hipLaunchKernelGGL(k_ndft_direct, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, nullptr, X, Y, u, v,
                       reinterpret_cast<const double2 *>(in), reinterpret_cast<double2 *>(out), 1.0 / Nu, 1.0 / Nv, M,
                       N);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// ============================================================================
// Part 2: the type-2 direct sum as a split-precision MFMA contraction
// (kernels in ndft2d_mfma.hpp).  A plan holds the mode operand.
namespace {

constexpr int kNdftMaxN = 8192;
constexpr int kNdftC = ndft::kChunkTiles;  // MFMA column tiles per chunk

size_t ndft_slice_bytes(int prec)
{
    switch (prec) {
    case NFFT2D_NDFT_FP32: return sizeof(ndft::Slice<ndft::kNdftFp32>);
    case NFFT2D_NDFT_BF16X3: return sizeof(ndft::Slice<ndft::kNdftBf16x3>);
    default: return sizeof(ndft::Slice<ndft::kNdftFp16>);
    }
}

}  // namespace

struct nfft2d_ndft_plan_s {
    int N1, N2, prec, device;
    int nch, nks;
    void *G;              // mode operand, nch * nks * C * 64 slices
    float *scale;         // N1 per-row scales and their inverses
    float *inv_scale;
};

extern "C" int nfft2d_ndft_plan_create(nfft2d_ndft_plan **plan, int N1, int N2, int precision)
{
    if (!plan) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_create: null plan pointer");
    *plan = nullptr;
    if (N1 <= 0 || N2 <= 0 || (N1 & 1) || (N2 & 1) || N1 > kNdftMaxN || N2 > kNdftMaxN)
        return nfft2d_fail(NFFT2D_ERR_SIZE, "nfft2d_ndft_plan_create: N1=%d N2=%d must be even, in [2, %d]", N1, N2,
                           kNdftMaxN);
    if (precision != NFFT2D_NDFT_FP32 && precision != NFFT2D_NDFT_BF16X3 && precision != NFFT2D_NDFT_FP16)
        return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_create: unknown precision %d", precision);
    nfft2d_ndft_plan_s *p = new (std::nothrow) nfft2d_ndft_plan_s();
    if (!p) return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft2d_ndft_plan_create: host allocation failed");
    p->N1 = N1;
    p->N2 = N2;
    p->prec = precision;
    NFFT2D_HIP_TRY(hipGetDevice(&p->device));
    p->nch = (N1 + ndft::kChunkK1 - 1) / ndft::kChunkK1;
    p->nks = (N2 + 8 * ndft::kKsPad - 1) / (8 * ndft::kKsPad) * ndft::kKsPad;
    const size_t gbytes = (size_t)p->nch * p->nks * kNdftC * 64 * ndft_slice_bytes(precision);
    if (hipMalloc(&p->G, gbytes) != hipSuccess || hipMalloc((void **)&p->scale, sizeof(float) * N1) != hipSuccess ||
        hipMalloc((void **)&p->inv_scale, sizeof(float) * N1) != hipSuccess) {
        (void)hipGetLastError();
        nfft2d_ndft_plan_destroy(p);
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft2d_ndft_plan_create: device allocation of %zu bytes failed", gbytes);
    }
    *plan = p;
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_plan_destroy(nfft2d_ndft_plan *p)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_plan_destroy: null plan");
    if (p->G) (void)hipFree(p->G);
    if (p->scale) (void)hipFree(p->scale);
    if (p->inv_scale) (void)hipFree(p->inv_scale);
    delete p;
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_set_modes(nfft2d_ndft_plan *p, const nfft2d_c64 *data, void *stream)
{
    if (!p || !data) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_set_modes: null argument");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const float2 *F = reinterpret_cast<const float2 *>(data);
    hipLaunchKernelGGL(ndft::k_ndft_rowscale, dim3(p->N1), dim3(256), 0, s, F, p->scale, p->inv_scale, p->N2,
                       p->prec == NFFT2D_NDFT_FP16 ? 1 : 0);
    NFFT2D_LAUNCH_CHECK();
    const long long total = (long long)p->nch * p->nks * kNdftC * 64;
    const dim3 grid((unsigned)((total + 255) / 256));
    switch (p->prec) {
    case NFFT2D_NDFT_FP32:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftFp32, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftFp32> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    case NFFT2D_NDFT_BF16X3:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftBf16x3, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftBf16x3> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    default:
        hipLaunchKernelGGL((ndft::k_ndft_prep<ndft::kNdftFp16, kNdftC>), grid, dim3(256), 0, s, F, p->scale,
                           (ndft::Slice<ndft::kNdftFp16> *)p->G, p->N1, p->N2, p->nch, p->nks);
        break;
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_eval(nfft2d_ndft_plan *p, nfft2d_c64 *result, const float *x, const float *y, int M,
                                void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: null plan");
    if (M < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: negative M");
    if (M == 0) return NFFT2D_OK;
    if (!result || !x || !y) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft2d_ndft_eval: null pointer");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((M + ndft::kPtsPerGroup - 1) / ndft::kPtsPerGroup));
    float2 *out = reinterpret_cast<float2 *>(result);
#ifndef NDFT_SPLIT
#define NDFT_SPLIT 2
#endif
    const dim3 block(64 * ndft::kWaves * NDFT_SPLIT);
    switch (p->prec) {
    case NFFT2D_NDFT_FP32:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftFp32, 4, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftFp32> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    case NFFT2D_NDFT_BF16X3:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftBf16x3, 16, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftBf16x3> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    default:
        hipLaunchKernelGGL((ndft::k_ndft_t2<ndft::kNdftFp16, 16, NDFT_SPLIT>), grid, block, 0, s,
                           (const ndft::Slice<ndft::kNdftFp16> *)p->G, p->inv_scale, x, y, out, p->N1, p->N2, M,
                           p->nch, p->nks);
        break;
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_ndft_type2(nfft2d_ndft_plan *p, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x,
                                 const float *y, int M, void *stream)
{
    int rc = nfft2d_ndft_set_modes(p, data, stream);
    if (rc != NFFT2D_OK) return rc;
    return nfft2d_ndft_eval(p, result, x, y, M, stream);
}
