// nfft2d_internal.hpp -- host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

// Records the calling thread's last error (read by nfft2d_last_error) and returns `code`.
int nfft2d_fail(int code, const char *fmt, ...);

#define NFFT2D_HIP_TRY(expr)                                                                              \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define NFFT2D_LAUNCH_CHECK()                                                                              \
    do {                                                                                                   \
        hipError_t e_ = hipGetLastError();                                                                 \
        if (e_ != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
    } while (0)
