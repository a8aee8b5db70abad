/*
 * nfft2d_overloads.hpp -- the reference's C++ call syntax over the C ABI of libnfft2d.so.
 *
 * The reference declares NFFT2_2D_GPU as two C++ overloads (float2 / double2,
 * NFFT/NFFT2_2D.cuh:8-11) and NDFT2_2D_GPU with a cublasHandle_t first
 * argument (NFFT/NDFT2_2D.cuh:4-6).  libnfft2d.so keeps its ABI free of C++
 * name mangling (NFFT2_2D_GPU_c64 / _c128, NDFT2_2D_GPU(void *, ...)); a
 * caller of the reference includes this header in place of "NFFT2_2D.cuh" /
 * "NDFT2_2D.cuh" and its calls compile unchanged, e.g.
 * Synthesis/PSO/Synthesis.cu:395
 *     NFFT2_2D_GPU(reinterpret_cast<float2 *>(d_far_field),
 *                  reinterpret_cast<float2 *>(d_excitations),
 *                  d_X_scaled, d_Y_scaled, 2 * Nv, 2 * Nu, M_x * M_y);
 * The reference's helper macros print and abort on an error (gpuErrchk,
 * cublasSafeCall); nfft2d_check does the same with the library's status and
 * message.  The plan warm-up calls (Calculate_cuFFT_plan_C2C_NFFT2_2D, ...)
 * are plain C in nfft2d.h already.
 */
#ifndef NFFT2D_OVERLOADS_HPP
#define NFFT2D_OVERLOADS_HPP

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "nfft2d.h"

inline void nfft2d_check(int rc, const char *what)
{
    if (rc != NFFT2D_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, nfft2d_last_error());
        std::exit(EXIT_FAILURE);
    }
}

/* NFFT/NFFT2_2D.cuh:8 (float path, NFFT/NFFT2_2D.cu:329). */
inline void NFFT2_2D_GPU(float2 *result, const float2 *data, const float *x, const float *y, const int N1,
                         const int N2, const int M)
{
    nfft2d_check(NFFT2_2D_GPU_c64(reinterpret_cast<nfft2d_c64 *>(result), reinterpret_cast<const nfft2d_c64 *>(data),
                                  x, y, N1, N2, M),
                 "NFFT2_2D_GPU");
}

/* NFFT/NFFT2_2D.cuh:10 (double path, NFFT/NFFT2_2D.cu:413). */
inline void NFFT2_2D_GPU(double2 *result, const double2 *data, const double *x, const double *y, const int N1,
                         const int N2, const int M)
{
    nfft2d_check(NFFT2_2D_GPU_c128(reinterpret_cast<nfft2d_c128 *>(result),
                                   reinterpret_cast<const nfft2d_c128 *>(data), x, y, N1, N2, M),
                 "NFFT2_2D_GPU");
}

/* NFFT/NDFT2_2D.cuh:4-6: the handle (a cublasHandle_t there) is accepted and
 * ignored; C is any 16-byte interleaved complex type (the reference's
 * double2_). */
template <typename H, typename C>
inline void NDFT2_2D_GPU(H handle, const double *X, const double *Y, const double *u, const double *v, C *in, C *out,
                         const int Nu, const int Nv, const int M, const int N)
{
    static_assert(sizeof(C) == sizeof(nfft2d_c128), "NDFT2_2D_GPU takes 16-byte complex values");
    nfft2d_check(::NDFT2_2D_GPU((void *)handle, X, Y, u, v, reinterpret_cast<const nfft2d_c128 *>(in),
                                reinterpret_cast<nfft2d_c128 *>(out), Nu, Nv, M, N),
                 "NDFT2_2D_GPU");
}

#endif /* NFFT2D_OVERLOADS_HPP */
