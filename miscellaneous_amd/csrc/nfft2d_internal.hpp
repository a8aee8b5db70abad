// nfft2d_internal.hpp -- host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

// Largest point count of one transform: keeps every point index and every
// per-block offset of the setpts kernels inside a signed 32-bit int.
constexpr int kMaxPoints = 1 << 30;

// Records the calling thread's last error (read by nfft2d_last_error) and returns `code`.
int nfft2d_fail(int code, const char *fmt, ...);

// NFFT2D_OK when `device` is the calling thread's current device, else an
// NFFT2D_ERR_ARG failure naming both.
int nfft2d_check_device(int device);

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

// Every device allocation of the library goes through here.  A test hook
// (nfft2d_debug_fail_alloc) can make one chosen allocation fail, so that each
// error path runs.
hipError_t nfft2d_malloc_raw(void **ptr, size_t bytes);
template <typename T>
inline hipError_t nfft2d_malloc(T **ptr, size_t bytes)
{
    return nfft2d_malloc_raw(reinterpret_cast<void **>(ptr), bytes);
}
