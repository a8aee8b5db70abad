// ndft2d.hip -- direct 2D NDFT in FP64, the reference's NDFT2_2D_GPU
// (NFFT/NDFT2_2D.cu:50-85):
//   out[n] = sum_{m<M} in[m] exp(-2 pi i (u[n] X[m] / Nu + v[n] Y[m] / Nv)).
// The reference fills an explicit M x N FP64 kernel matrix (:57-63; 64 GiB at
// M = N = 65536), writes it to a text file (:69) and calls cublasZgemv (:80).
// Here the kernel is evaluated on the fly: each thread owns one output n and
// streams the points through LDS in chunks, accumulating in FP64.
#include <hip/hip_runtime.h>

#include "../../include/nfft2d.h"
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
