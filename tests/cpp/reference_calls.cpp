// reference_calls.cpp -- the reference's own native calls, compiled against
// include/nfft2d_overloads.hpp and linked with libnfft2d.so.
//
// Reproduces the far-field step of Synthesis/PSO/Synthesis.cu:360-420 at its
// real size: Nu = Nv = 44, an M_x x M_y = 44 x 44 element lattice (M = 1936)
// scaled to grid units, and the calls, with the reference's argument order,
//   Calculate_cuFFT_plan_C2C_NFFT2_2D(2 * Nu, 2 * Nv);                 PSO.cu:216
//   NFFT2_2D_GPU(float2 *, float2 *, d_X_scaled, d_Y_scaled, 2 * Nv, 2 * Nu, M_x * M_y);   Synthesis.cu:395
//   Calculate_cuFFT_plan_Z2Z_NFFT2_2D(2 * Nu, 2 * Nv);                 Synthesis.cu:506
//   NFFT2_2D_GPU(double2 *, double2 *, ..., 2 * Nv, 2 * Nu, M_x * M_y); Synthesis.cu:508
//   NDFT2_2D_GPU(handle, d_X_scaled, d_Y_scaled, d_U_discrete, d_V_discrete, d_excitations,
//                d_far_field, 2 * Nu, 2 * Nv, M_x * M_y, (2 * Nu) * (2 * Nv));      Synthesis.cu:397,510
// with (U, V) the integer meshgrid of Synthesis.cu:85-90 (u, v = -Nu .. Nu-1).
// Inputs and outputs go to the binary file named by argv[1]; the test that
// runs this program checks them against the CPU oracle.  With --type3 it
// makes the type-3 call of the reference's MEX gateway instead (below).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "nfft2d_overloads.hpp"

#define HIP_OK(e)                                                                               \
    do {                                                                                        \
        hipError_t r_ = (e);                                                                    \
        if (r_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "%s: %s\n", #e, hipGetErrorString(r_));                       \
            return 2;                                                                           \
        }                                                                                       \
    } while (0)

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double urand()   // [0, 1), splitmix64
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T> static T *dev_copy(const std::vector<T> &h)
{
    T *d = nullptr;
    if (hipMalloc(&d, sizeof(T) * (h.size() ? h.size() : 1)) != hipSuccess) return nullptr;
    if (!h.empty() && hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

// The type-3 call of NFFT/NFFT3/C/NFFT3_2D_C_mex.cpp:18-65 with NUFFT_2D_CPU
// replaced by NUFFT_2D_GPU_host (same arguments; fftw_complex is double[2]):
// host arrays x, y (len_in), s, t (len_out), f as interleaved (re, im) built
// from separate real and imaginary parts (:37-46), eps = 1e-9.  Inputs and
// outputs go to `path`.
static int type3_mex_call(const char *path)
{
    typedef double fftw_complex[2];
    const int len_in = 3000, len_out = 2000;
    std::vector<double> x(len_in), y(len_in), s(len_out), t(len_out), fr(len_in), fi(len_in);
    for (int i = 0; i < len_in; ++i) {
        x[i] = -6 + 12 * urand();
        y[i] = -2 + 7 * urand();
        fr[i] = urand() - 0.5;
        fi[i] = urand() - 0.5;
    }
    for (int i = 0; i < len_out; ++i) {
        s[i] = -9 + 14 * urand();
        t[i] = -4 + 8 * urand();
    }
    std::vector<fftw_complex> f(len_in), F(len_out);
    for (int i = 0; i < len_in; i++) {
        f[i][0] = fr[i];
        f[i][1] = fi[i];
    }
    const double eps[1] = {1e-9};
    const int rc = NUFFT_2D_GPU_host(x.data(), y.data(), s.data(), t.data(), (const nfft2d_c128 *)f.data(),
                                     (nfft2d_c128 *)F.data(), eps[0], len_in, len_out);
    if (rc != NFFT2D_OK) {
        std::fprintf(stderr, "NUFFT_2D_GPU_host: %s\n", nfft2d_last_error());
        return 2;
    }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return 1;
    const int32_t hdr[2] = {len_in, len_out};
    std::fwrite(hdr, sizeof(hdr), 1, fp);
    std::fwrite(x.data(), sizeof(double), len_in, fp);
    std::fwrite(y.data(), sizeof(double), len_in, fp);
    std::fwrite(s.data(), sizeof(double), len_out, fp);
    std::fwrite(t.data(), sizeof(double), len_out, fp);
    std::fwrite(f.data(), sizeof(fftw_complex), len_in, fp);
    std::fwrite(F.data(), sizeof(fftw_complex), len_out, fp);
    std::fwrite(eps, sizeof(double), 1, fp);
    std::fclose(fp);
    std::printf("reference_calls: NUFFT_2D_GPU_host (the NUFFT_2D_CPU call of NFFT3_2D_C_mex.cpp:65), %d -> %d: done\n",
                len_in, len_out);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "--type3") return type3_mex_call(argv[2]);
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <output file> | --type3 <output file>\n", argv[0]);
        return 1;
    }
    const int Nu = 44, Nv = 44, M_x = 44, M_y = 44, M = M_x * M_y, N = (2 * Nu) * (2 * Nv);
    // element lattice (spacing ~1.9 grid units, jittered) in grid units, as the
    // caller's scaling by 2 Nu (pi / a') / (2 pi) leaves it; rounded to float
    // once so that the float and double calls see the same coordinates
    std::vector<float> Xf(M), Yf(M);
    std::vector<double> Xd(M), Yd(M);
    std::vector<float2> ef(M);
    std::vector<double2> ed(M);
    for (int iy = 0; iy < M_y; ++iy)
        for (int ix = 0; ix < M_x; ++ix) {
            const int m = iy * M_x + ix;
            Xf[m] = (float)(-41.5 + 1.9 * ix + 0.6 * (urand() - 0.5));
            Yf[m] = (float)(-41.5 + 1.9 * iy + 0.6 * (urand() - 0.5));
            Xd[m] = Xf[m];
            Yd[m] = Yf[m];
            const double a = 2.0 * 3.141592653589793 * urand(), r = 0.5 + urand();
            ed[m] = make_double2(r * std::cos(a), r * std::sin(a));
            ef[m] = make_float2((float)ed[m].x, (float)ed[m].y);
            ed[m] = make_double2(ef[m].x, ef[m].y);
        }
    // the integer meshgrid of Synthesis.cu:85-90, row-major: U along the first output index
    std::vector<double> U(N), V(N);
    for (int i = 0; i < 2 * Nu; ++i)
        for (int j = 0; j < 2 * Nv; ++j) {
            U[i * (2 * Nv) + j] = -Nu + i;
            V[i * (2 * Nv) + j] = -Nv + j;
        }
    float *d_Xf = dev_copy(Xf), *d_Yf = dev_copy(Yf);
    double *d_Xd = dev_copy(Xd), *d_Yd = dev_copy(Yd), *d_U = dev_copy(U), *d_V = dev_copy(V);
    float2 *d_ef = dev_copy(ef);
    double2 *d_ed = dev_copy(ed);
    float2 *d_ff = nullptr;
    double2 *d_fd = nullptr, *d_fn = nullptr;
    HIP_OK(hipMalloc(&d_ff, sizeof(float2) * N));
    HIP_OK(hipMalloc(&d_fd, sizeof(double2) * N));
    HIP_OK(hipMalloc(&d_fn, sizeof(double2) * N));
    if (!d_Xf || !d_Yf || !d_Xd || !d_Yd || !d_U || !d_V || !d_ef || !d_ed) {
        std::fprintf(stderr, "device allocation failed\n");
        return 2;
    }
    void *cublasHandleSynthesis = nullptr;   // the reference's handle, accepted and ignored

    Calculate_cuFFT_plan_C2C_NFFT2_2D(2 * Nu, 2 * Nv);
    NFFT2_2D_GPU(reinterpret_cast<float2 *>(d_ff), reinterpret_cast<float2 *>(d_ef), d_Xf, d_Yf, 2 * Nv, 2 * Nu, M_x * M_y);
    Calculate_cuFFT_plan_Z2Z_NFFT2_2D(2 * Nu, 2 * Nv);
    NFFT2_2D_GPU(reinterpret_cast<double2 *>(d_fd), reinterpret_cast<double2 *>(d_ed), d_Xd, d_Yd, 2 * Nv, 2 * Nu,
                 M_x * M_y);
    NDFT2_2D_GPU(cublasHandleSynthesis, d_Xd, d_Yd, d_U, d_V, d_ed, d_fn, 2 * Nu, 2 * Nv, M_x * M_y,
                 (2 * Nu) * (2 * Nv));
    HIP_OK(hipDeviceSynchronize());
    Destroy_cuFFT_plan_NFFT2_2D();

    std::vector<float2> ff(N);
    std::vector<double2> fd(N), fn(N);
    HIP_OK(hipMemcpy(ff.data(), d_ff, sizeof(float2) * N, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(fd.data(), d_fd, sizeof(double2) * N, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(fn.data(), d_fn, sizeof(double2) * N, hipMemcpyDeviceToHost));
    FILE *fp = std::fopen(argv[1], "wb");
    if (!fp) {
        std::fprintf(stderr, "cannot write %s\n", argv[1]);
        return 1;
    }
    const int32_t hdr[4] = {Nu, Nv, M, N};
    std::fwrite(hdr, sizeof(hdr), 1, fp);
    std::fwrite(Xd.data(), sizeof(double), M, fp);
    std::fwrite(Yd.data(), sizeof(double), M, fp);
    std::fwrite(ed.data(), sizeof(double2), M, fp);
    std::fwrite(ff.data(), sizeof(float2), N, fp);
    std::fwrite(fd.data(), sizeof(double2), N, fp);
    std::fwrite(fn.data(), sizeof(double2), N, fp);
    std::fclose(fp);
    hipFree(d_Xf); hipFree(d_Yf); hipFree(d_Xd); hipFree(d_Yd); hipFree(d_U); hipFree(d_V);
    hipFree(d_ef); hipFree(d_ed); hipFree(d_ff); hipFree(d_fd); hipFree(d_fn);
    std::printf("reference_calls: NFFT2_2D_GPU float2 / double2 and NDFT2_2D_GPU at 88 x 88, M = %d: done\n", M);
    return 0;
}
