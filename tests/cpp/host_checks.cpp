// host_checks.cpp -- the library's host code under AddressSanitizer.
//
// Built from the library sources with the host side instrumented
// (hipcc -Xarch_host -fsanitize=address; device code is not), linked into this
// program, and run in one of two modes:
//   args  every entry point with invalid arguments (null pointers, negative or
//         oversized counts, wrong dtypes, destroyed-handle patterns): each must
//         return a negative status and set a message; runs with or without a
//         GPU (without one, calls that reach the device report NFFT2D_ERR_HIP);
//   full  (needs the GPU) small transforms through every family of entry
//         points, then the same sequence again with each device allocation in
//         turn made to fail (nfft2d_debug_fail_alloc): every failure must be
//         reported, objects must be destroyable afterwards, and a clean rerun
//         must give the same results.
// Any out-of-bounds access, use after free or double free in the host code
// stops the program with an AddressSanitizer report (non-zero exit status).
// Also a regression check for the reference's read of x[i], y[i] before its
// i < M guard (NFFT/NFFT2_2D.cu:167,171): the host code here reads nothing
// before its checks, which the args mode exercises with M = 0 and null arrays.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "nfft2d.h"

static int g_fail = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);            \
            std::fprintf(stderr, "\n");                   \
            ++g_fail;                                     \
        }                                                 \
    } while (0)

static void expect_err(int rc, const char *what)
{
    CHECK(rc < 0, "%s returned %d, expected an error", what);
    CHECK(rc >= 0 || std::strlen(nfft2d_last_error()) > 0, "%s: no error message", what);
}

static uint64_t g_state = 0x2545F4914F6CDD1Dull;
static double urand()   // [0, 1), splitmix64
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------ args mode
static void check_args()
{
    nfft2d_plan *p = nullptr;
    expect_err(nfft2d_plan_create(nullptr, 64, 64, NFFT2D_C64, 6, 2, 0), "plan_create(null out)");
    expect_err(nfft2d_plan_create(&p, 63, 64, NFFT2D_C64, 6, 2, 0), "plan_create(odd N1)");
    expect_err(nfft2d_plan_create(&p, 64, -2, NFFT2D_C64, 6, 2, 0), "plan_create(N2 < 0)");
    expect_err(nfft2d_plan_create(&p, 64, 64, 7, 6, 2, 0), "plan_create(bad dtype)");
    expect_err(nfft2d_plan_create(&p, 64, 64, NFFT2D_C64, 5, 2, 0), "plan_create(K = 5)");
    expect_err(nfft2d_plan_create(&p, 64, 64, NFFT2D_C64, 6, 3, 0), "plan_create(c = 3)");
    expect_err(nfft2d_plan_create(&p, 64, 64, NFFT2D_C64, 6, 2, 1u), "plan_create(flags)");
    CHECK(p == nullptr, "plan_create set the handle on failure");
    expect_err(nfft2d_plan_destroy(nullptr), "plan_destroy(null)");
    expect_err(nfft2d_setpts(nullptr, nullptr, nullptr, 0, nullptr), "setpts(null plan)");
    expect_err(nfft2d_type2(nullptr, nullptr, nullptr, nullptr), "type2(null plan)");
    expect_err(nfft2d_type1(nullptr, nullptr, nullptr, nullptr), "type1(null plan)");
    expect_err(nfft2d_plan_num_points(nullptr), "plan_num_points(null)");
    expect_err(NFFT2_2D_GPU_c64(nullptr, nullptr, nullptr, nullptr, 64, 64, 10), "NFFT2_2D_GPU_c64(null result)");
    expect_err(NFFT2_2D_GPU_c64((nfft2d_c64 *)8, nullptr, nullptr, nullptr, 64, 64, 10),
               "NFFT2_2D_GPU_c64(null data, M > 0)");
    expect_err(NFFT2_2D_GPU_c64((nfft2d_c64 *)8, nullptr, nullptr, nullptr, 64, 64, -1), "NFFT2_2D_GPU_c64(M < 0)");
    expect_err(NFFT1_2D_GPU_c128(nullptr, nullptr, nullptr, nullptr, 64, 64, 10), "NFFT1_2D_GPU_c128(null)");
    expect_err(NFFT2_2D_GPU_c64_noalloc((nfft2d_c64 *)8, (nfft2d_c64 *)8, (float *)8, (float *)8, 512, 512, 10,
                                        nullptr),
               "noalloc(N above its limit)");
    expect_err(NFFT2_2D_GPU_c64_noalloc((nfft2d_c64 *)8, (nfft2d_c64 *)8, (float *)8, (float *)8, 64, 64,
                                        1 << 20, nullptr),
               "noalloc(M above its limit)");
    expect_err(NDFT2_2D_GPU(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 8, 8, 10, 10),
               "NDFT2_2D_GPU(null arrays)");
    expect_err(NDFT2_2D_GPU(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 8, 8, -1, 10),
               "NDFT2_2D_GPU(M < 0)");
    nfft2d_ndft_plan *q = nullptr;
    expect_err(nfft2d_ndft_plan_create(nullptr, 64, 64, 0), "ndft_plan_create(null out)");
    expect_err(nfft2d_ndft_plan_create(&q, 0, 64, 0), "ndft_plan_create(N1 = 0)");
    expect_err(nfft2d_ndft_plan_create(&q, 64, 64, 9), "ndft_plan_create(bad precision)");
    expect_err(nfft2d_ndft_plan_destroy(nullptr), "ndft_plan_destroy(null)");
    nfft1d_plan *l = nullptr;
    expect_err(nfft1d_plan_create(nullptr, 64), "nfft1d_plan_create(null out)");
    expect_err(nfft1d_plan_create(&l, 63), "nfft1d_plan_create(odd N)");
    expect_err(nfft1d_type2(nullptr, nullptr, nullptr, nullptr, 10, 1, nullptr), "nfft1d_type2(null plan)");
    nfft2d_batch *b = nullptr;
    int n1[2] = {64, 63}, n2[2] = {64, 64}, mm[2] = {10, 10};
    expect_err(nfft2d_batch_create(&b, 2, n1, n2, mm, 0), "batch_create(odd N)");
    expect_err(nfft2d_batch_create(&b, -1, n1, n2, mm, 0), "batch_create(count < 0)");
    expect_err(nfft2d_batch_create(nullptr, 1, n1, n2, mm, 0), "batch_create(null out)");
    expect_err(nfft2d_batch_create(&b, 1, nullptr, n2, mm, 0), "batch_create(null sizes)");
    expect_err(nfft2d_batch_destroy(nullptr), "batch_destroy(null)");
    expect_err(nfft2d_batch_type2(nullptr, nullptr), "batch_type2(null)");
    expect_err(NUFFT_2D_GPU(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1e-6, 10, 10), "NUFFT_2D_GPU(null)");
    expect_err(NUFFT_2D_GPU((double *)8, (double *)8, (double *)8, (double *)8, (nfft2d_c128 *)8, (nfft2d_c128 *)8,
                            0.0, 10, 10),
               "NUFFT_2D_GPU(eps = 0)");
    expect_err(NUFFT_2D_GPU_host(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1e-6, -1, 10),
               "NUFFT_2D_GPU_host(len_in < 0)");
    expect_err(NUFFT_2D_GPU_params(nullptr, nullptr, nullptr, nullptr, 1e-6, 10, 10, nullptr),
               "NUFFT_2D_GPU_params(null)");
    expect_err(nfft2d_batch_set_direct(nullptr, 0, 8), "batch_set_direct(null)");
    expect_err(nfft2d_batch_set_gather_blocks(nullptr, 1024), "batch_set_gather_blocks(null)");
    const double *dp = reinterpret_cast<const double *>(8);
    expect_err(nfft2d_knab_stats(nullptr, 8, 1, dp, 8, 1, dp, dp, 4, 4, nullptr, (double *)8, nullptr),
               "knab_stats(null axis)");
    expect_err(nfft2d_knab_stats(dp, 1, 1, dp, 8, 1, dp, dp, 4, 4, nullptr, (double *)8, nullptr),
               "knab_stats(one-sample axis)");
    expect_err(nfft2d_knab_spectrum(NFFT2D_C64, dp, 8, 8, 5, 2, nullptr, nullptr, 1.0, (void *)8, 8, 8, nullptr),
               "knab_spectrum(band wider than the padding)");
    expect_err(nfft2d_knab_spectrum(NFFT2D_C64, dp, 8, 8, 1, 1, dp, nullptr, 1.0, (void *)8, 8, 8, nullptr),
               "knab_spectrum(one window table)");
    expect_err(nfft2d_knab_points(7, dp, dp, 4, 4, nullptr, 4, 4, 0.0, 0.0, 1.0, 1.0, 8.0, 8.0, (void *)8, (void *)8,
                                  nullptr),
               "knab_points(bad dtype)");
    expect_err(nfft2d_knab_points(NFFT2D_C64, dp, dp, 4, 4, nullptr, 3, 4, 0.0, 0.0, 1.0, 1.0, 8.0, 8.0, (void *)8,
                                  (void *)8, nullptr),
               "knab_points(padded grid smaller)");
    expect_err(nfft2d_knab_finish(NFFT2D_C128, dp, 4, 4, 3, 1.0, dp, dp, nullptr, 0, 1, 0, 1, 0, 0, nullptr, nullptr),
               "knab_finish(null result)");
}

// ------------------------------------------------------------------ full mode (GPU)
struct Buf {
    void *d = nullptr;
    size_t n = 0;
    explicit Buf(size_t bytes) : n(bytes)
    {
        if (hipMalloc(&d, bytes) != hipSuccess) d = nullptr;
    }
    ~Buf() { (void)hipFree(d); }
    template <typename T>
    T *as() const { return static_cast<T *>(d); }
    void up(const void *h) const { (void)hipMemcpy(d, h, n, hipMemcpyHostToDevice); }
    std::vector<unsigned char> down() const
    {
        std::vector<unsigned char> h(n);
        (void)hipMemcpy(h.data(), d, n, hipMemcpyDeviceToHost);
        return h;
    }
};

template <typename T>
static std::vector<T> rnd(size_t n, double lo, double hi)
{
    std::vector<T> v(n);
    for (auto &e : v) e = (T)(lo + (hi - lo) * urand());
    return v;
}

// One scenario: returns 0 or the first error status; its outputs are appended to `out`.
using Scenario = std::function<int(std::vector<std::vector<unsigned char>> &out)>;

static int plan_scenario(std::vector<std::vector<unsigned char>> &out)
{
    const int N = 96, M = 5000;
    g_state = 11;
    auto x = rnd<float>(M, -N / 2, N / 2), y = rnd<float>(M, -N / 2, N / 2);
    auto f = rnd<float>(2 * N * N, -1, 1), d = rnd<float>(2 * M, -1, 1);
    Buf dx(4 * M), dy(4 * M), df(8 * N * N), dd(8 * M), r2(8 * M), r1(8 * N * N);
    dx.up(x.data()); dy.up(y.data()); df.up(f.data()); dd.up(d.data());
    nfft2d_plan *p = nullptr;
    int rc = nfft2d_plan_create(&p, N, N, NFFT2D_C64, 6, 2, 0);
    if (rc) return rc;
    if (!(rc = nfft2d_setpts(p, dx.d, dy.d, M, nullptr)) && !(rc = nfft2d_type2(p, r2.d, df.d, nullptr)) &&
        !(rc = nfft2d_type1(p, r1.d, dd.d, nullptr))) {
        (void)hipDeviceSynchronize();
        out.push_back(r2.down());
        out.push_back(r1.down());
    }
    CHECK(nfft2d_plan_destroy(p) == 0, "plan_destroy after status %d", rc);
    return rc;
}

static int batch_scenario(std::vector<std::vector<unsigned char>> &out)
{
    g_state = 12;
    int N1[2] = {64, 128}, N2[2] = {64, 96}, M[2] = {3000, 4000};
    std::vector<Buf *> bufs;
    const void *xs[2], *ys[2], *fs[2], *ds[2];
    void *o2[2], *o1[2];
    for (int t = 0; t < 2; ++t) {
        auto x = rnd<float>(M[t], -N1[t] / 2, N1[t] / 2), y = rnd<float>(M[t], -N2[t] / 2, N2[t] / 2);
        auto f = rnd<float>(2 * N1[t] * N2[t], -1, 1), d = rnd<float>(2 * M[t], -1, 1);
        Buf *bx = new Buf(4 * M[t]), *by = new Buf(4 * M[t]), *bf = new Buf(8 * N1[t] * N2[t]),
            *bd = new Buf(8 * M[t]), *b2 = new Buf(8 * M[t]), *b1 = new Buf(8 * N1[t] * N2[t]);
        bx->up(x.data()); by->up(y.data()); bf->up(f.data()); bd->up(d.data());
        bufs.insert(bufs.end(), {bx, by, bf, bd, b2, b1});
        xs[t] = bx->d; ys[t] = by->d; fs[t] = bf->d; ds[t] = bd->d; o2[t] = b2->d; o1[t] = b1->d;
    }
    nfft2d_batch *b = nullptr;
    int rc = nfft2d_batch_create(&b, 2, N1, N2, M, 0);
    if (!rc) {
        if (!(rc = nfft2d_batch_set_points(b, (const float *const *)xs, (const float *const *)ys)) &&
            !(rc = nfft2d_batch_set_io(b, (const nfft2d_c64 *const *)fs, (nfft2d_c64 *const *)o2)) &&
            !(rc = nfft2d_batch_setpts_type2(b, nullptr)) &&
            !(rc = nfft2d_batch_set_io_type1(b, (const nfft2d_c64 *const *)ds, (nfft2d_c64 *const *)o1)) &&
            !(rc = nfft2d_batch_type1(b, nullptr))) {
            (void)hipDeviceSynchronize();
            for (int i = 0; i < 2; ++i) {
                out.push_back(bufs[6 * i + 4]->down());
                out.push_back(bufs[6 * i + 5]->down());
            }
        }
        CHECK(nfft2d_batch_destroy(b) == 0, "batch_destroy after status %d", rc);
    }
    for (Buf *q : bufs) delete q;
    return rc;
}

static int ndft_scenario(std::vector<std::vector<unsigned char>> &out)
{
    g_state = 13;
    const int N = 32, M = 2000;
    auto f = rnd<float>(2 * N * N, -1, 1), x = rnd<float>(M, -N / 2, N / 2), y = rnd<float>(M, -N / 2, N / 2);
    Buf df(8 * N * N), dx(4 * M), dy(4 * M), o(8 * M);
    df.up(f.data()); dx.up(x.data()); dy.up(y.data());
    nfft2d_ndft_plan *q = nullptr;
    int rc = nfft2d_ndft_plan_create(&q, N, N, 0);
    if (rc) return rc;
    if (!(rc = nfft2d_ndft_set_modes(q, df.as<nfft2d_c64>(), nullptr)) &&
        !(rc = nfft2d_ndft_eval(q, o.as<nfft2d_c64>(), dx.as<float>(), dy.as<float>(), M, nullptr))) {
        (void)hipDeviceSynchronize();
        out.push_back(o.down());
    }
    CHECK(nfft2d_ndft_plan_destroy(q) == 0, "ndft_plan_destroy after status %d", rc);
    return rc;
}

static int plan1d_scenario(std::vector<std::vector<unsigned char>> &out)
{
    g_state = 14;
    const int N = 64, M = 3000, B = 2;
    auto f = rnd<float>(2 * B * N, -1, 1), x = rnd<float>(B * M, -N / 2, N / 2);
    Buf df(8 * B * N), dx(4 * B * M), o(8 * B * M);
    df.up(f.data()); dx.up(x.data());
    nfft1d_plan *l = nullptr;
    int rc = nfft1d_plan_create(&l, N);
    if (rc) return rc;
    if (!(rc = nfft1d_type2(l, o.as<nfft2d_c64>(), df.as<nfft2d_c64>(), dx.as<float>(), M, B, nullptr))) {
        (void)hipDeviceSynchronize();
        out.push_back(o.down());
    }
    CHECK(nfft1d_plan_destroy(l) == 0, "nfft1d_plan_destroy after status %d", rc);
    return rc;
}

static int dropin_scenario(std::vector<std::vector<unsigned char>> &out)
{
    g_state = 15;
    const int N = 44, M = 1000;
    auto x = rnd<float>(M, -N / 2, N / 2), y = rnd<float>(M, -N / 2, N / 2), d = rnd<float>(2 * M, -1, 1);
    Buf dx(4 * M), dy(4 * M), dd(8 * M), o(8 * N * N);
    dx.up(x.data()); dy.up(y.data()); dd.up(d.data());
    const int rc = NFFT2_2D_GPU_c64(o.as<nfft2d_c64>(), dd.as<nfft2d_c64>(), dx.as<float>(), dy.as<float>(), N, N, M);
    if (!rc) {
        (void)hipDeviceSynchronize();
        out.push_back(o.down());
    }
    Destroy_cuFFT_plan_NFFT2_2D();   // no cached plan: every run allocates
    return rc;
}

static int type3_scenario(std::vector<std::vector<unsigned char>> &out)
{
    g_state = 16;
    const int n = 1500;
    auto x = rnd<double>(n, -3, 3), y = rnd<double>(n, -2, 2), s = rnd<double>(n, -5, 5),
         t = rnd<double>(n, -4, 4), f = rnd<double>(2 * n, -1, 1);
    std::vector<double> F(2 * n);
    const int rc = NUFFT_2D_GPU_host(x.data(), y.data(), s.data(), t.data(), (const nfft2d_c128 *)f.data(),
                                     (nfft2d_c128 *)F.data(), 1e-9, n, n);
    if (!rc) {
        std::vector<unsigned char> b(F.size() * 8);
        std::memcpy(b.data(), F.data(), b.size());
        out.push_back(b);
    }
    return rc;
}

static bool near(const std::vector<std::vector<unsigned char>> &a, const std::vector<std::vector<unsigned char>> &b,
                 bool dbl)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        if (a[i].size() != b[i].size()) return false;
        const size_t n = a[i].size() / (dbl ? 8 : 4);
        double num = 0, den = 0;
        for (size_t k = 0; k < n; ++k) {
            const double u = dbl ? ((const double *)a[i].data())[k] : ((const float *)a[i].data())[k];
            const double v = dbl ? ((const double *)b[i].data())[k] : ((const float *)b[i].data())[k];
            if (!std::isfinite(u) || !std::isfinite(v)) return false;
            num += (u - v) * (u - v);
            den += v * v;
        }
        if (num > 1e-10 * den) return false;   // type-1 sums follow arrival order inside a tile unit
    }
    return true;
}

static void check_full()
{
    struct Named {
        const char *name;
        Scenario run;
        bool dbl;
    } scs[] = {{"plan", plan_scenario, false},     {"batch", batch_scenario, false},
               {"ndft", ndft_scenario, false},     {"plan1d", plan1d_scenario, false},
               {"dropin", dropin_scenario, false}, {"type3", type3_scenario, true}};
    for (auto &sc : scs) {
        std::vector<std::vector<unsigned char>> ref, warm;
        nfft2d_debug_fail_alloc(-1);
        CHECK(sc.run(warm) == 0, "%s: warm-up failed: %s", sc.name, nfft2d_last_error());
        const long long c0 = nfft2d_debug_alloc_count();
        CHECK(sc.run(ref) == 0, "%s: clean run failed: %s", sc.name, nfft2d_last_error());
        const long long n = nfft2d_debug_alloc_count() - c0;
        std::printf("%s: %lld device allocations per run\n", sc.name, n);
        for (long long i = 0; i < n; ++i) {
            std::vector<std::vector<unsigned char>> junk, again;
            nfft2d_debug_fail_alloc(i);
            const int rc = sc.run(junk);
            nfft2d_debug_fail_alloc(-1);
            CHECK(rc == NFFT2D_ERR_ALLOC, "%s: allocation %lld failed but the status was %d (%s)", sc.name, i, rc,
                  nfft2d_last_error());
            CHECK(sc.run(again) == 0 && near(again, ref, sc.dbl), "%s: clean rerun after failing allocation %lld",
                  sc.name, i);
        }
    }
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "args";
    check_args();
    if (mode == "full") check_full();
    std::printf("%s: %s (%d failures)\n", mode.c_str(), g_fail ? "FAILED" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
