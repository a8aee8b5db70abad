// nfft2d_api.hip -- host side of libnfft2d.so: plans, FP64 tables, launches,
// and the C ABI declared in include/nfft2d.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>
#include <algorithm>
#include <atomic>

#include "../../include/nfft2d.h"
#include "nfft2d_internal.hpp"
#include "nfft2d_kernels.hpp"
#include "nfft1d_kernels.hpp"
#include "nfft2d_czt.hpp"

using namespace nfft2d;

// ============================================================== errors
static thread_local std::string g_last_error;

int nfft2d_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" const char *nfft2d_last_error(void) { return g_last_error.c_str(); }

// ============================================================== device allocations
static std::atomic<long long> g_alloc_seq{0};     // allocations attempted so far
static std::atomic<long long> g_alloc_fail{-1};   // sequence number of the one to fail, or -1

hipError_t nfft2d_malloc_raw(void **ptr, size_t bytes)
{
    const long long k = g_alloc_seq.fetch_add(1);
    if (k == g_alloc_fail.load()) {
        *ptr = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(ptr, bytes);
}

extern "C" long long nfft2d_debug_alloc_count(void) { return g_alloc_seq.load(); }
extern "C" void nfft2d_debug_fail_alloc(long long n) { g_alloc_fail.store(n < 0 ? -1 : g_alloc_seq.load() + n); }
extern "C" int nfft2d_version(void) { return 100; }

// ============================================================== FP64 window math (host)
static constexpr double kPi = 3.141592653589793238462643383279502884;
static double host_alpha() { return (2.0 - 1.0 / kC) * kPi - 0.01; }   // NFFT/NFFT2_2D.cu:27

// I0 by its power series sum_j (x^2/4)^j / (j!)^2.
static double host_i0(double x)
{
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int j = 1; j < 1000; ++j) {
        term *= q / ((double)j * (double)j);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

// f(P) = sinh(alpha sqrt P)/sqrt P continued through P <= 0, and f'(P).
static double knab_f(double P, double a)
{
    if (std::fabs(P) < 1.0) {
        double a2 = a * a, term = a, sum = a;
        for (int j = 1; j < 60; ++j) {
            term *= a2 * P / ((2.0 * j) * (2.0 * j + 1.0));
            sum += term;
        }
        return sum;
    }
    if (P > 0) { double s = std::sqrt(P); return std::sinh(a * s) / s; }
    double s = std::sqrt(-P);
    return std::sin(a * s) / s;
}
static double knab_fprime(double P, double a)
{
    if (std::fabs(P) < 1.0) {
        double a2 = a * a, c = a * a2 / 6.0, Pp = 1.0, sum = c;
        for (int j = 2; j < 60; ++j) {
            c *= a2 / ((2.0 * j) * (2.0 * j + 1.0));
            Pp *= P;
            sum += (double)j * c * Pp;
        }
        return sum;
    }
    if (P > 0) { double s = std::sqrt(P); return (a * s * std::cosh(a * s) - std::sinh(a * s)) / (2 * s * s * s); }
    double s = std::sqrt(-P);
    return -(a * s * std::cos(a * s) - std::sin(a * s)) / (2 * s * s * s);
}
// phi_hat(d) and phi_hat'(d) (NFFT/NFFT2_2D.cu:112-122).
static double host_phihat(double d) { return knab_f(kK * kK - d * d, host_alpha()) / kPi; }
static double host_phihat_d(double d) { return -2.0 * d * knab_fprime(kK * kK - d * d, host_alpha()) / kPi; }

// Interpolating polynomial of degree Deg at the Chebyshev nodes of
// [-1/2, 1/2], monomial coefficients, solved in long double.
template <int Deg, typename F>
static void fit_poly(F f, double *coef)
{
    constexpr int n = Deg + 1;
    long double A[n][n + 1];
    for (int i = 0; i < n; ++i) {
        long double t = 0.5L * std::cos(kPi * (i + 0.5) / n);
        long double p = 1;
        for (int j = 0; j < n; ++j) { A[i][j] = p; p *= t; }
        A[i][n] = f((double)t);
    }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs((double)A[r][c]) > std::fabs((double)A[piv][c])) piv = r;
        for (int k = 0; k <= n; ++k) std::swap(A[c][k], A[piv][k]);
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            long double m = A[r][c] / A[c][c];
            for (int k = c; k <= n; ++k) A[r][k] -= m * A[c][k];
        }
    }
    for (int i = 0; i < n; ++i) coef[i] = (double)(A[i][n] / A[i][i]);
}

static void make_window_args(WinArgs &wa)
{
    const double S = host_i0(kK * host_alpha());
    std::memset(&wa, 0, sizeof(wa));
    double cf[kPolyDeg + 1];
    for (int j = 0; j <= kK; ++j) {
        fit_poly<kPolyDeg>([&](double t) { return host_phihat(t - j) / S; }, cf);
        for (int q = 0; q < kNE; ++q) wa.val.e[j][q] = (float)cf[2 * q];
        for (int q = 0; q < kNO; ++q) wa.val.o[j][q] = j == 0 ? 0.f : (float)cf[2 * q + 1];
        fit_poly<kPolyDeg>([&](double t) { return host_phihat_d(t - j) / S; }, cf);
        for (int q = 0; q < kNE; ++q) wa.der.e[j][q] = j == 0 ? 0.f : (float)cf[2 * q];
        for (int q = 0; q < kNO; ++q) wa.der.o[j][q] = (float)cf[2 * q + 1];
    }
}

// The FP64 tables: the same fits at degree kPolyDeg64.
static void make_window_tables64(WinTables64 &wt)
{
    const double S = host_i0(kK * host_alpha());
    std::memset(&wt, 0, sizeof(wt));
    double cf[kPolyDeg64 + 1];
    for (int j = 0; j <= kK; ++j) {
        fit_poly<kPolyDeg64>([&](double t) { return host_phihat(t - j) / S; }, cf);
        for (int q = 0; q < kNE64; ++q) wt.val.e[j][q] = cf[2 * q];
        for (int q = 0; q < kNO64; ++q) wt.val.o[j][q] = j == 0 ? 0.0 : cf[2 * q + 1];
        fit_poly<kPolyDeg64>([&](double t) { return host_phihat_d(t - j) / S; }, cf);
        for (int q = 0; q < kNE64; ++q) wt.der.e[j][q] = j == 0 ? 0.0 : cf[2 * q];
        for (int q = 0; q < kNO64; ++q) wt.der.o[j][q] = cf[2 * q + 1];
    }
}

// ============================================================== plan
// A second stream and two events, for running a plan's (or batch's) bin sort
// beside its grid passes: fork = record on the caller's stream s, aux waits;
// join = record on aux, s waits.  Both are stream-ordered, so the pattern is
// captured into a hipGraph like any other work on s (as two branches).  The
// stream and the events are created at the first use (not inside a capture:
// the first fused call also sizes the point buffers, as nfft2d_setpts does).
struct Fork {
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int init(int device)
    {
        if (aux) return NFFT2D_OK;
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(device) != hipSuccess)
            return nfft2d_fail(NFFT2D_ERR_HIP, "fork: cannot select the device");
        const bool ok = hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) == hipSuccess &&
                        hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) == hipSuccess &&
                        hipEventCreateWithFlags(&ev_join, hipEventDisableTiming) == hipSuccess;
        (void)hipSetDevice(cur);
        if (!ok) {
            destroy();
            return nfft2d_fail(NFFT2D_ERR_HIP, "fork: cannot create the second stream");
        }
        return NFFT2D_OK;
    }
    int fork(hipStream_t s)
    {
        if (hipEventRecord(ev_fork, s) != hipSuccess || hipStreamWaitEvent(aux, ev_fork, 0) != hipSuccess)
            return nfft2d_fail(NFFT2D_ERR_HIP, "fork: event record / wait failed");
        return NFFT2D_OK;
    }
    int join(hipStream_t s)
    {
        if (hipEventRecord(ev_join, aux) != hipSuccess || hipStreamWaitEvent(s, ev_join, 0) != hipSuccess)
            return nfft2d_fail(NFFT2D_ERR_HIP, "join: event record / wait failed");
        return NFFT2D_OK;
    }
    void destroy()
    {
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux) (void)hipStreamDestroy(aux);
        aux = nullptr;
        ev_fork = ev_join = nullptr;
    }
};

struct nfft2d_plan_s {
    int N1 = 0, N2 = 0, L1 = 0, L2 = 0, dtype = 0, device = 0;
    int nb1 = 0, nb2 = 0, nbins = 0, maxpts = kGatherThreads;
    size_t esz = 0;                       // sizeof(complex)
    void *tw1 = nullptr, *tw2 = nullptr;  // 2N1 / 2N2 twiddles e^{-2 pi i q/(2N)}
    void *D1 = nullptr, *D2 = nullptr;    // deconvolution S / phi(xi_k)
    void *grid = nullptr;                 // L2 x L1, transposed oversampled grid
    void *tmp = nullptr;                  // N1 x L2 intermediate
    int M = -1, Mcap = 0, subcap = 0, histcap = 0;
    int nblk = 0, chunk = 0;              // setpts blocking of the last point set
    void *pts = nullptr;                  // sorted point records PtRec<T>
    int *idx = nullptr, *rank = nullptr, *hist = nullptr, *tot = nullptr, *nsub = nullptr;
    int4 *sub = nullptr;
    int *subbeg = nullptr;                // first subproblem of each bin (nbins + 1)
    int2 *gexcl = nullptr, *gsum = nullptr;  // bin-scan prefixes inside / of groups of kColBins bins
    void *w64 = nullptr;                  // FP64 plans: WinTables64 in device memory
    void *scratch = nullptr;              // MFMA spread tiles, 28x28 per subproblem
    int scratchcap = 0;
    void *resid = nullptr;                // residual of the last nfft2d_type2_lsq, sorted point order
    int residcap = 0;
    bool resid_valid = false;             // resid belongs to the current point set
    BinGeom geo;
    WinArgs wa;
    Fork fk;                              // setpts beside the grid passes (nfft2d_setpts_type2)
    // chirp-z tables of a dimension whose N is not a power of two (null otherwise):
    // [cin2 N][cout2 2N][hhat2 m][cin1 2N][cout1 N][hhat1 m][T1 m/64][T2 64]
    void *czt[2] = {nullptr, nullptr};
    int cztm[2] = {0, 0};
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
static int max_fft_n(int dtype) { return dtype == NFFT2D_C64 ? 4096 : 2048; }
static bool fast_path(int N, int dtype) { return is_pow2(N) && N >= 8 && N <= max_fft_n(dtype); }
// Lengths that are not powers of two take the chirp-z passes from 128 on (below, the
// O(N^2) direct passes cost less than two m-point FFTs: the PSO generation's 250
// transforms of 88 x 88 took 0.97 ms through chirp-z passes, 0.89 through direct ones)
static constexpr int kCztMinN = 128;
// Largest chirp-z line: its m-point FFT (m >= 3N - 1) in LDS, at most 16384 complex64 / 8192 complex128
template <typename T> constexpr int kCztMaxN = sizeof(T) == 4 ? 4096 : 2048;
static bool czt_path(int N, int dtype) { return !fast_path(N, dtype) && N >= kCztMinN; }

static void hfree(void *ptr)
{
    if (ptr) (void)hipFree(ptr);
}

static void free_points(nfft2d_plan *p)
{
    hfree(p->pts); hfree(p->rank); hfree(p->sub);
    p->pts = nullptr; p->rank = nullptr; p->sub = nullptr;
    p->Mcap = 0;
    p->subcap = 0;
}

static void free_scratch(nfft2d_plan *p)
{
    hfree(p->scratch);
    p->scratch = nullptr;
    p->scratchcap = 0;
}

static void free_hist(nfft2d_plan *p)
{
    hfree(p->hist);
    p->hist = nullptr;
    p->histcap = 0;
}

static void free_plan(nfft2d_plan *p)
{
    if (!p) return;
    for (void *&c : p->czt) {
        hfree(c);
        c = nullptr;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(p->device);
    hfree(p->tw1); hfree(p->tw2); hfree(p->D1); hfree(p->D2);
    hfree(p->grid); hfree(p->tmp); hfree(p->w64);
    hfree(p->tot); hfree(p->nsub);
    free_points(p);
    free_hist(p);
    free_scratch(p);
    hfree(p->resid);
    p->fk.destroy();
    hfree(p->subbeg);
    hfree(p->gexcl); hfree(p->gsum);
    (void)hipSetDevice(cur);
    delete p;
}

// In-place radix-2 FFT on the host, e^{-2 pi i jk / n}, n a power of two.
static void host_fft(std::vector<std::complex<double>> &a)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const double ang = -2.0 * kPi * (double)k / (double)len;
                const std::complex<double> w(std::cos(ang), std::sin(ang));
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// The chirp-z tables of one dimension (N not a power of two), in FP64 on the
// host: c[k] = e^{-i pi k^2 / L}, L = 2N (k^2 reduced mod 2L exactly); type-2
// passes p0 = -N/2, q0 = 0 (P = N, Q = L), type-1 passes p0 = 0, q0 = -N/2
// (P = L, Q = N); the kernel conj(c[d + q0 - p0]) at d mod m, d in [1 - P, Q),
// transformed; the FFT's two twiddle tables.
template <typename T>
static int czt_buffer(int N, void **out, int *mout)
{
    using C = cplx<T>;
    const int L = 2 * N, m = czt_m(N), nt1 = m >= 64 ? m / 64 : 1;
    auto chirp = [&](long long k) {
        const long long q = ((k * k) % (2LL * L) + 2LL * L) % (2LL * L);
        const double ang = -kPi * (double)q / (double)L;
        return std::complex<double>(std::cos(ang), std::sin(ang));
    };
    std::vector<C> h;
    h.reserve((size_t)3 * L + 2 * (size_t)m + nt1 + 64);
    auto put = [&](std::complex<double> v) {
        C c;
        c.x = (T)v.real();
        c.y = (T)v.imag();
        h.push_back(c);
    };
    auto pass = [&](int P, int Q, int p0, int q0) {
        for (int i = 0; i < P; ++i) put(chirp(i + p0));
        for (int i = 0; i < Q; ++i) put(chirp(i + q0));
        std::vector<std::complex<double>> ker(m, 0.0);
        for (int d = 1 - P; d < Q; ++d) ker[(size_t)((d % m) + m) % m] = std::conj(chirp((long long)d + q0 - p0));
        host_fft(ker);
        for (int i = 0; i < m; ++i) put(ker[i]);
    };
    pass(N, L, -N / 2, 0);   // type-2
    pass(L, N, 0, -N / 2);   // type-1
    for (int j = 0; j < nt1; ++j) put(std::polar(1.0, -2.0 * kPi * 64.0 * j / m));
    for (int j = 0; j < 64; ++j) put(std::polar(1.0, -2.0 * kPi * (double)j / m));
    if (nfft2d_malloc(out, sizeof(C) * h.size()) != hipSuccess)
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "plan tables: hipMalloc of the chirp-z tables failed");
    NFFT2D_HIP_TRY(hipMemcpy(*out, h.data(), sizeof(C) * h.size(), hipMemcpyHostToDevice));
    *mout = m;
    return NFFT2D_OK;
}
template <typename T>
static int upload_czt(nfft2d_plan *p, int dim, int N)
{
    return czt_buffer<T>(N, &p->czt[dim], &p->cztm[dim]);
}

template <typename T>
static CztTables<T> czt_tables(const nfft2d_plan *p, int dim, bool t2)
{
    return czt_tables_of<T>(p->czt[dim], dim == 0 ? p->N1 : p->N2, p->cztm[dim], t2);
}

// Launches one chirp-z pass: `lines` lines (one block each).
template <typename T, int SIGN, bool SCALE_IN, int U>
static int czt_launch_u(const CztTables<T> &tb, int lines, const cplx<T> *in, long long in_ls, long long in_es,
                        cplx<T> *out, long long out_ls, long long out_es, const T *sl, const T *se, hipStream_t s)
{
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute((const void *)k_czt<T, SIGN, SCALE_IN, U>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(cplx<T>) * czt_lds_entries(czt_m(kCztMaxN<T>))));
    });
    if (attr != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "chirp-z pass: %s", hipGetErrorString(attr));
    const int nt = std::min(1024, std::max(64, tb.m / 4));
    hipLaunchKernelGGL((k_czt<T, SIGN, SCALE_IN, U>), dim3(lines), dim3(nt), sizeof(cplx<T>) * czt_lds_entries(tb.m),
                       s, in, in_ls, in_es, out, out_ls, out_es, sl, se, tb);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}
template <typename T, int SIGN, bool SCALE_IN>
static int czt_launch(const CztTables<T> &tb, int lines, const cplx<T> *in, long long in_ls, long long in_es,
                      cplx<T> *out, long long out_ls, long long out_es, const T *sl, const T *se, hipStream_t s)
{
    // at 1024 threads a thread holds 8 values up to m = 8192, 16 at m = 16384
    if (tb.m > 8192)
        return czt_launch_u<T, SIGN, SCALE_IN, 4>(tb, lines, in, in_ls, in_es, out, out_ls, out_es, sl, se, s);
    return czt_launch_u<T, SIGN, SCALE_IN, 2>(tb, lines, in, in_ls, in_es, out, out_ls, out_es, sl, se, s);
}

template <typename T>
static int upload_tables(nfft2d_plan *p)
{
    using C = cplx<T>;
    const double S = host_i0(kK * host_alpha());
    const double a = host_alpha();
    for (int dim = 0; dim < 2; ++dim) {
        const int N = dim == 0 ? p->N1 : p->N2;
        std::vector<C> tw(2 * N);
        for (int q = 0; q < 2 * N; ++q) {
            double ang = -kPi * (double)q / (double)N;   // -2 pi q / (2N)
            tw[q].x = (T)std::cos(ang);
            tw[q].y = (T)std::sin(ang);
        }
        // D[k] = S / I0(K sqrt(alpha^2 - xi_k^2)), xi_k = 2 pi k / (c N)  (NFFT/NFFT2_2D.cu:60-69)
        std::vector<T> D(N);
        for (int i = 0; i < N; ++i) {
            double xi = 2.0 * kPi * (double)(i - N / 2) / ((double)kC * N);
            D[i] = (T)(S / host_i0(kK * std::sqrt(a * a - xi * xi)));
        }
        void *&twd = dim == 0 ? p->tw1 : p->tw2;
        void *&Dd = dim == 0 ? p->D1 : p->D2;
        if (nfft2d_malloc(&twd, sizeof(C) * 2 * N) != hipSuccess || nfft2d_malloc(&Dd, sizeof(T) * N) != hipSuccess)
            return nfft2d_fail(NFFT2D_ERR_ALLOC, "plan tables: hipMalloc failed");
        NFFT2D_HIP_TRY(hipMemcpy(twd, tw.data(), sizeof(C) * 2 * N, hipMemcpyHostToDevice));
        NFFT2D_HIP_TRY(hipMemcpy(Dd, D.data(), sizeof(T) * N, hipMemcpyHostToDevice));
        if (czt_path(N, p->dtype)) {
            int rc = upload_czt<T>(p, dim, N);
            if (rc) return rc;
        }
    }
    return NFFT2D_OK;
}

extern "C" int nfft2d_plan_create(nfft2d_plan **out, int N1, int N2, int dtype, int K, int c, unsigned flags)
{
    if (!out) return nfft2d_fail(NFFT2D_ERR_ARG, "plan_create: null output pointer");
    *out = nullptr;
    if (dtype != NFFT2D_C64 && dtype != NFFT2D_C128) return nfft2d_fail(NFFT2D_ERR_ARG, "plan_create: bad dtype %d", dtype);
    if (flags != 0) return nfft2d_fail(NFFT2D_ERR_ARG, "plan_create: flags must be 0");
    if (K != kK || c != kC) return nfft2d_fail(NFFT2D_ERR_UNSUPPORTED, "plan_create: only K=6, c=2 are supported (got K=%d, c=%d)", K, c);
    const int maxN = dtype == NFFT2D_C64 ? 4096 : 2048;
    if (N1 <= 0 || N2 <= 0 || (N1 & 1) || (N2 & 1) || N1 > maxN || N2 > maxN)
        return nfft2d_fail(NFFT2D_ERR_SIZE, "plan_create: N1, N2 must be even and in [2, %d] (got %d, %d)", maxN, N1, N2);
    static_assert(kCztMaxN<float> >= 4096 && kCztMaxN<double> >= 2048, "every even N up to the maximum has passes");
    int device = 0;
    NFFT2D_HIP_TRY(hipGetDevice(&device));      // before the allocation: nothing to free on failure
    nfft2d_plan *p = new (std::nothrow) nfft2d_plan_s();
    if (!p) return nfft2d_fail(NFFT2D_ERR_ALLOC, "plan_create: host allocation failed");
    p->N1 = N1; p->N2 = N2; p->L1 = kC * N1; p->L2 = kC * N2; p->dtype = dtype;
    p->esz = dtype == NFFT2D_C64 ? sizeof(float2) : sizeof(double2);
    p->device = device;
    p->nb1 = (p->L1 + kBin - 1) / kBin;
    p->nb2 = (p->L2 + kBin - 1) / kBin;
    p->nbins = p->nb1 * p->nb2;
    p->geo.L1 = p->L1;
    p->geo.L2 = p->L2;
    p->geo.nb1 = p->nb1;
    p->geo.inv1d = 1.0 / p->L1;
    p->geo.inv2d = 1.0 / p->L2;
    p->geo.inv1f = (float)p->geo.inv1d;
    p->geo.inv2f = (float)p->geo.inv2d;
    make_window_args(p->wa);
    int rc = dtype == NFFT2D_C64 ? upload_tables<float>(p) : upload_tables<double>(p);
    if (rc) { free_plan(p); return rc; }
    if (dtype == NFFT2D_C128) {
        WinTables64 wt;
        make_window_tables64(wt);
        if (nfft2d_malloc(&p->w64, sizeof(wt)) != hipSuccess) {
            free_plan(p);
            return nfft2d_fail(NFFT2D_ERR_ALLOC, "plan_create: hipMalloc of the FP64 window tables failed");
        }
        if (hipMemcpy(p->w64, &wt, sizeof(wt), hipMemcpyHostToDevice) != hipSuccess) {
            free_plan(p);
            return nfft2d_fail(NFFT2D_ERR_HIP, "plan_create: upload of the FP64 window tables failed");
        }
        p->wa.w64 = (const WinTables64 *)p->w64;
    }
    if (nfft2d_malloc(&p->grid, p->esz * (size_t)p->L1 * p->L2) != hipSuccess ||
        nfft2d_malloc(&p->tmp, p->esz * (size_t)p->N1 * p->L2) != hipSuccess ||
        nfft2d_malloc(&p->tot, sizeof(int) * p->nbins) != hipSuccess ||
        nfft2d_malloc(&p->nsub, sizeof(int)) != hipSuccess ||
        nfft2d_malloc(&p->subbeg, sizeof(int) * (p->nbins + 1)) != hipSuccess ||
        nfft2d_malloc(&p->gexcl, sizeof(int2) * p->nbins) != hipSuccess ||
        nfft2d_malloc(&p->gsum, sizeof(int2) * ((p->nbins + kColBins - 1) / kColBins)) != hipSuccess) {
        free_plan(p);
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "plan_create: hipMalloc of the grid failed");
    }
    *out = p;
    return NFFT2D_OK;
}

extern "C" int nfft2d_plan_destroy(nfft2d_plan *p)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "plan_destroy: null plan");
    free_plan(p);
    return NFFT2D_OK;
}

extern "C" int nfft2d_plan_set_maxpts(nfft2d_plan *p, int maxpts)
{
    if (!p || maxpts < 32 || maxpts > kGatherThreads)
        return nfft2d_fail(NFFT2D_ERR_ARG, "set_maxpts: maxpts must be in [32, %d]", kGatherThreads);
    p->maxpts = maxpts;
    p->M = -1;   // subproblems must be rebuilt
    p->resid_valid = false;
    return NFFT2D_OK;
}

extern "C" int nfft2d_plan_num_points(const nfft2d_plan *p)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "null plan");
    if (p->M < 0) return nfft2d_fail(NFFT2D_ERR_STATE, "plan has no points");
    return p->M;
}

// A plan's buffers live on the device that was current at its creation; a
// call made with another device current would launch there on foreign
// pointers, so it is refused.
int nfft2d_check_device(int device)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "hipGetDevice failed");
    if (cur != device)
        return nfft2d_fail(NFFT2D_ERR_ARG, "plan belongs to device %d but device %d is current", device, cur);
    return NFFT2D_OK;
}

static unsigned grid1d(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T>
static int zero_fill(void *ptr, long long n, hipStream_t s)
{
    if (n <= 0) return NFFT2D_OK;
    unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_zero<T>, dim3(blocks), dim3(256), 0, s, (T *)ptr, n);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// Blocking of setpts pass 1 for M points: LDS histograms for up to
// kMaxLdsBins bins (nblk blocks of chunk points), else one logical block with
// global atomics.
static void setpts_blocking(const nfft2d_plan *p, int M, int &nblk, int &chunk)
{
    const bool lds = p->nbins <= kMaxLdsBins;
    nblk = lds ? (int)std::max<long long>(1, std::min<long long>(256, ((long long)M + 2047) / 2048)) : 1;
    chunk = std::max(1, (M + nblk - 1) / nblk);
}

// Point-dependent buffers of the plan for M points (grown, never shrunk).
template <typename T>
static int ensure_point_buffers(nfft2d_plan *p, int M)
{
    const int subcap = p->nbins + (M + p->maxpts - 1) / p->maxpts;
    if (M > p->Mcap || subcap > p->subcap) {
        free_points(p);
        const int cap = std::max(M, 1);
        if (nfft2d_malloc(&p->pts, sizeof(PtRec<T>) * cap) != hipSuccess ||
            nfft2d_malloc(&p->rank, sizeof(int) * cap) != hipSuccess ||
            nfft2d_malloc(&p->sub, sizeof(int4) * std::max(subcap, 1)) != hipSuccess) {
            free_points(p);
            p->M = -1;
            return nfft2d_fail(NFFT2D_ERR_ALLOC, "setpts: hipMalloc for %d points failed", M);
        }
        p->Mcap = cap;
        p->subcap = std::max(subcap, 1);
    }
    int nblk, chunk;
    setpts_blocking(p, M, nblk, chunk);
    const long long hneed = (long long)nblk * p->nbins;
    if (hneed > p->histcap) {
        free_hist(p);
        if (nfft2d_malloc(&p->hist, sizeof(int) * hneed) != hipSuccess) {
            p->M = -1;
            return nfft2d_fail(NFFT2D_ERR_ALLOC, "setpts: hipMalloc of the bin histogram failed");
        }
        p->histcap = (int)hneed;
    }
    return NFFT2D_OK;
}

template <typename T>
static int setpts_impl(nfft2d_plan *p, const T *x, const T *y, int M, hipStream_t s)
{
    int rc = ensure_point_buffers<T>(p, M);
    if (rc) return rc;
    int nblk, chunk;
    setpts_blocking(p, M, nblk, chunk);
    const bool lds = p->nbins <= kMaxLdsBins;
    const T inv1 = (T)(1.0 / p->L1), inv2 = (T)(1.0 / p->L2);
    if (lds) {
        hipLaunchKernelGGL(k_hist<T>, dim3(nblk), dim3(kHistThreads), sizeof(int) * p->nbins, s, x, y, M, chunk, p->L1,
                           p->L2, inv1, inv2, p->nb1, p->nbins, p->hist, (int *)nullptr);
    } else {
        int rc = zero_fill<int>(p->hist, p->nbins, s);
        if (rc) return rc;
        if (M > 0)
            hipLaunchKernelGGL(k_hist_global<T>, dim3(grid1d(M, 256)), dim3(256), 0, s, x, y, M, p->L1, p->L2, inv1,
                               inv2, p->nb1, p->hist, p->rank);
    }
    NFFT2D_LAUNCH_CHECK();
    const int ngroups = (int)grid1d(p->nbins, kColBins);
    hipLaunchKernelGGL(k_colscan, dim3(ngroups), dim3(256), 0, s, p->hist, nblk, p->nbins, p->maxpts, p->tot, p->gexcl,
                       p->gsum);
    NFFT2D_LAUNCH_CHECK();
    if (lds) {
        // ranks recomputed by the scatter blocks, which take the histogram blocks' points
        hipLaunchKernelGGL(k_scatter_lds<T>, dim3(nblk), dim3(kScatThreads), 0, s, x, y, M, chunk, p->L1, p->L2, inv1,
                           inv2, p->nb1, p->nbins, p->maxpts, p->hist, p->tot, p->gexcl, p->gsum, ngroups,
                           (PtRec<T> *)p->pts, p->subbeg, p->sub, p->nsub);
        NFFT2D_LAUNCH_CHECK();
    } else if (ngroups <= kScatThreads) {
        const long long nb = std::max(grid1d(M, kScatThreads * kScatPer), grid1d(p->nbins, kScatThreads));
        hipLaunchKernelGGL(k_scatter_scan<T>, dim3(nb), dim3(kScatThreads), 0, s, x, y, M, chunk, p->L1, p->L2, inv1,
                           inv2, p->nb1, p->nbins, p->maxpts, p->hist, p->tot, p->gexcl, p->gsum, ngroups, p->rank,
                           (PtRec<T> *)p->pts, p->subbeg, p->sub, p->nsub);
        NFFT2D_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_scan_groups, dim3(1), dim3(kScanThreads), 0, s, p->gsum, ngroups, p->nbins, p->subbeg,
                           p->nsub);
        NFFT2D_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_scatter<T>, dim3(grid1d(std::max(M, p->nbins), 256)), dim3(256), 0, s, x, y, M, chunk,
                           p->L1, p->L2, inv1, inv2, p->nb1, p->nbins, p->maxpts, p->hist, p->tot, p->gexcl, p->gsum,
                           p->rank, (PtRec<T> *)p->pts, p->subbeg, p->sub);
        NFFT2D_LAUNCH_CHECK();
    }
    p->nblk = nblk;
    p->chunk = chunk;
    p->M = M;
    p->resid_valid = false;
    return NFFT2D_OK;
}

extern "C" int nfft2d_setpts(nfft2d_plan *p, const void *x, const void *y, int M, void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "setpts: null plan");
    if (M < 0 || M > kMaxPoints) return nfft2d_fail(NFFT2D_ERR_ARG, "setpts: M = %d outside [0, %d]", M, kMaxPoints);
    if (M > 0 && (!x || !y)) return nfft2d_fail(NFFT2D_ERR_ARG, "setpts: null coordinates");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    return p->dtype == NFFT2D_C64 ? setpts_impl<float>(p, (const float *)x, (const float *)y, M, s)
                                  : setpts_impl<double>(p, (const double *)x, (const double *)y, M, s);
}

// ============================================================== pass dispatch
template <typename F>
static bool dispatch_n(int N, F &&f)
{
    switch (N) {
    case 8: f(std::integral_constant<int, 8>{}); return true;
    case 16: f(std::integral_constant<int, 16>{}); return true;
    case 32: f(std::integral_constant<int, 32>{}); return true;
    case 64: f(std::integral_constant<int, 64>{}); return true;
    case 128: f(std::integral_constant<int, 128>{}); return true;
    case 256: f(std::integral_constant<int, 256>{}); return true;
    case 512: f(std::integral_constant<int, 512>{}); return true;
    case 1024: f(std::integral_constant<int, 1024>{}); return true;
    case 2048: f(std::integral_constant<int, 2048>{}); return true;
    case 4096: f(std::integral_constant<int, 4096>{}); return true;
    default: return false;
    }
}

template <typename T> static constexpr int max_n() { return sizeof(T) == 4 ? 4096 : 2048; }

// Type-2 passes (a)+(b): data (N1 x N2) -> transposed oversampled grid.
template <typename T>
static int t2_grid(nfft2d_plan *p, const void *data, hipStream_t s)
{
    using C = cplx<T>;
    const C *f = (const C *)data;
    C *Tm = (C *)p->tmp, *G = (C *)p->grid;
    const T *D1 = (const T *)p->D1, *D2 = (const T *)p->D2;
    const C *tw1 = (const C *)p->tw1, *tw2 = (const C *)p->tw2;
    if (fast_path(p->N2, p->dtype)) {
        dispatch_n(p->N2, [&](auto n) {
            constexpr int N = decltype(n)::value;
            if constexpr (N <= max_n<T>()) {
                constexpr int R = rows_per_block<N>();
                hipLaunchKernelGGL((k_t2_rows<T, N, -1>), dim3(grid1d(p->N1, R)), dim3(R * (N / 8)), 0, s, f, Tm, D1,
                                   D2, tw2, p->N1);
            }
        });
    } else if (p->czt[1]) {   // rows: f[k1][*] (N2) -> Tm[k1][*] (L2), pre-compensated
        int rc = czt_launch<T, -1, true>(czt_tables<T>(p, 1, true), p->N1, f, p->N2, 1, Tm, p->L2, 1, D1, D2, s);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL((k_t2_rows_direct<T, -1>), dim3(p->N1), dim3(256), sizeof(C) * 3 * p->N2, s, f, Tm, D1, D2, tw2,
                           p->N1, p->N2);
    }
    NFFT2D_LAUNCH_CHECK();
    if (fast_path(p->N1, p->dtype)) {
        dispatch_n(p->N1, [&](auto n) {
            constexpr int N = decltype(n)::value;
            if constexpr (N <= max_n<T>()) {
                constexpr int B = cols_per_block<N>();
                hipLaunchKernelGGL((k_t2_cols<T, N, -1>), dim3(grid1d(p->L2, B)), dim3(B * (N / 8)), 0, s, Tm, G, tw1,
                                   p->L2);
            }
        });
    } else if (p->czt[0]) {   // columns: Tm[*][n2] (N1) -> G[n2][*] (L1)
        int rc = czt_launch<T, -1, true>(czt_tables<T>(p, 0, true), p->L2, Tm, 1, p->L2, G, p->L1, 1, (const T *)nullptr,
                                         (const T *)nullptr, s);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL((k_t2_cols_direct<T, -1>), dim3(p->L2), dim3(256), sizeof(C) * 3 * p->N1, s, Tm, G, tw1, p->N1,
                           p->L2);
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// complex64 lines of at most kT1MmMaxN modes that take neither the
// fixed-size FFT nor the chirp-z passes run on the matrix cores.
template <typename T>
static bool t1_plan_mfma(int N)
{
    return std::is_same<T, float>::value && N <= kT1MmMaxN && (N & 1) == 0;
}

// Type-1 passes: transposed grid -> crop + deconvolve -> result (N1 x N2).
template <typename T, int SIGN>
static int t1_passes(nfft2d_plan *p, void *result, hipStream_t s)
{
    using C = cplx<T>;
    C *Tp = (C *)p->tmp, *G = (C *)p->grid, *F = (C *)result;
    const T *D1 = (const T *)p->D1, *D2 = (const T *)p->D2;
    const C *tw1 = (const C *)p->tw1, *tw2 = (const C *)p->tw2;
    if (fast_path(p->N1, p->dtype)) {
        dispatch_n(p->N1, [&](auto n) {
            constexpr int N = decltype(n)::value;
            if constexpr (N <= max_n<T>()) {
                constexpr int B = cols_per_block<N>();
                hipLaunchKernelGGL((k_t1_cols<T, N, SIGN>), dim3(grid1d(p->L2, B)), dim3(B * (N / 8)), 0, s, G, Tp,
                                   tw1, p->L2);
            }
        });
    } else if (p->czt[0]) {   // columns: G[n2][*] (L1) -> Tp[*][n2] (N1)
        int rc = czt_launch<T, SIGN, false>(czt_tables<T>(p, 0, false), p->L2, G, p->L1, 1, Tp, 1, p->L2,
                                            (const T *)nullptr, (const T *)nullptr, s);
        if (rc) return rc;
    } else if (t1_plan_mfma<T>(p->N1)) {
        hipLaunchKernelGGL((k_t1_cols_mm<SIGN>), dim3(grid1d(p->L2, 16), grid1d(p->N1, 16)), dim3(256), 0, s, (const float2 *)G,
                           (float2 *)Tp, (const float2 *)tw1, p->N1, p->L2);
    } else {
        hipLaunchKernelGGL((k_t1_cols_direct<T, SIGN>), dim3(p->L2), dim3(256), sizeof(C) * 2 * p->L1, s, G, Tp, tw1, p->N1,
                           p->L2);
    }
    NFFT2D_LAUNCH_CHECK();
    if (fast_path(p->N2, p->dtype)) {
        dispatch_n(p->N2, [&](auto n) {
            constexpr int N = decltype(n)::value;
            if constexpr (N <= max_n<T>()) {
                constexpr int R = rows_per_block<N>();
                hipLaunchKernelGGL((k_t1_rows<T, N, SIGN>), dim3(grid1d(p->N1, R)), dim3(R * (N / 8)), 0, s, Tp, F, D1,
                                   D2, tw2, p->N1);
            }
        });
    } else if (p->czt[1]) {   // rows: Tp[k1][*] (L2) -> F[k1][*] (N2), deconvolved
        int rc = czt_launch<T, SIGN, false>(czt_tables<T>(p, 1, false), p->N1, Tp, p->L2, 1, F, p->N2, 1, D1, D2, s);
        if (rc) return rc;
    } else if (t1_plan_mfma<T>(p->N2)) {
        hipLaunchKernelGGL((k_t1_rows_mm<SIGN>), dim3(grid1d(p->N1, 16), grid1d(p->N2, 16)), dim3(256), 0, s, (const float2 *)Tp,
                           (float2 *)F, (const float *)D1, (const float *)D2, (const float2 *)tw2, p->N1, p->N2);
    } else {
        hipLaunchKernelGGL((k_t1_rows_direct<T, SIGN>), dim3(p->N1), dim3(256), sizeof(C) * 2 * p->L2, s, Tp, F, D1, D2,
                           tw2, p->N1, p->N2);
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// the subproblem capacity rounded up to a multiple of 64 (the XCD-ordered kernels deal runs of 8)
static unsigned sub_blocks(const nfft2d_plan *p) { return (unsigned)((p->subcap + 63) & ~63); }

template <typename T, int MODE>
static int gather(nfft2d_plan *p, void *out, void *d1, void *d2, const void *g, void *gx1, void *gx2, hipStream_t s)
{
    using C = cplx<T>;
    if (p->M == 0) return NFFT2D_OK;
    if constexpr (std::is_same<T, float>::value && MODE == 0) {
        // Small tails (a bin's last subproblem) exist where bins hold more than
        // maxpts points: from 128 points per bin on average the merged form is
        // used (1M polar points, 512^2 modes: 33.3 against 34.3 us; 4M uniform
        // points, 1024^2 modes: 130 against 144 us).  On sparse sets its
        // neighbour reads cost more than it saves (64K uniform points, 512^2
        // modes: 17.6 against 16.4 us).
        const bool dense = (long long)p->M >= 128LL * p->nbins;
        auto kern = dense ? k_gather_pairs<2> : k_gather_pairs<1>;
        hipLaunchKernelGGL(kern, dim3(sub_blocks(p)), dim3(kGatherThreads), 0, s, (const float2 *)p->grid, p->geo,
                           (const PtRec<float> *)p->pts, p->sub, p->nsub, p->wa, (float2 *)out);
        NFFT2D_LAUNCH_CHECK();
        return NFFT2D_OK;
    }
    hipLaunchKernelGGL((k_gather<T, MODE>), dim3(sub_blocks(p)), dim3(kGatherThreads), 0, s, (const C *)p->grid, p->geo,
                       (const PtRec<T> *)p->pts, p->sub, p->nsub, p->wa, (C *)out, (C *)d1, (C *)d2, (const C *)g,
                       (T *)gx1, (T *)gx2);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

static bool mfma_spread_ok(const nfft2d_plan *p)
{
    return p->nb1 >= 3 && p->nb2 >= 3 && p->L1 >= 44 && p->L2 >= 44;
}

// data: M samples in the callers' point order, or (sorted) in the plan's
// sorted point-record order.
template <typename T>
static int spread_impl(nfft2d_plan *p, const void *data, bool sorted, hipStream_t s)
{
    using C = cplx<T>;
    if (mfma_spread_ok(p)) {
        if (p->subcap > p->scratchcap) {
            free_scratch(p);
            if (nfft2d_malloc(&p->scratch, sizeof(C) * kTileCells * (size_t)p->subcap) != hipSuccess)
                return nfft2d_fail(NFFT2D_ERR_ALLOC, "spread: hipMalloc of %d tiles failed", p->subcap);
            p->scratchcap = p->subcap;
        }
        if (p->M > 0) {
            if constexpr (std::is_same<T, float>::value)
                hipLaunchKernelGGL((sorted ? k_spread_halo<true> : k_spread_halo<false>), dim3(sub_blocks(p)),
                                   dim3(256), 0, s, (const PtRec<float> *)p->pts, (const float2 *)data, p->sub,
                                   p->nsub, p->subbeg, p->tot, p->maxpts, p->geo, p->wa, (float2 *)p->scratch);
            else
                hipLaunchKernelGGL((sorted ? k_spread_halo64<true> : k_spread_halo64<false>), dim3(sub_blocks(p)),
                                   dim3(256), 0, s, (const PtRec<double> *)p->pts, (const double2 *)data, p->sub,
                                   p->nsub, p->subbeg, p->tot, p->maxpts, p->geo, p->wa, (double2 *)p->scratch);
            NFFT2D_LAUNCH_CHECK();
        }
        // a dimension whose last bin is thinner than K needs the +-2 neighbours
        const int r1 = merge_radius(p->L1), r2 = merge_radius(p->L2);
        auto merge = r1 == 1 ? (r2 == 1 ? k_spread_merge<T, 1, 1> : k_spread_merge<T, 1, 2>)
                             : (r2 == 1 ? k_spread_merge<T, 2, 1> : k_spread_merge<T, 2, 2>);
        hipLaunchKernelGGL(merge, dim3(p->nbins), dim3(256), 0, s, (const C *)p->scratch, p->subbeg, p->geo, p->nb2,
                           (C *)p->grid);
        NFFT2D_LAUNCH_CHECK();
        return NFFT2D_OK;
    }
    int rc = zero_fill<T>(p->grid, 2LL * p->L1 * p->L2, s);
    if (rc) return rc;
    if (p->M > 0) {
        hipLaunchKernelGGL((sorted ? k_spread<T, true> : k_spread<T, false>), dim3(sub_blocks(p)), dim3(256), 0, s,
                           (C *)p->grid, p->geo, (const PtRec<T> *)p->pts, (const C *)data, p->sub, p->nsub, p->wa);
        NFFT2D_LAUNCH_CHECK();
    }
    return NFFT2D_OK;
}

static int check_exec(const nfft2d_plan *p, const void *a, const void *b)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "null plan");
    if (p->M < 0) return nfft2d_fail(NFFT2D_ERR_STATE, "plan has no points: call nfft2d_setpts first");
    if (p->M > 0 && (!a || !b)) return nfft2d_fail(NFFT2D_ERR_ARG, "null data or result pointer");
    return nfft2d_check_device(p->device);
}

extern "C" int nfft2d_type2_grid(nfft2d_plan *p, const void *data, void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "null plan");
    if (!data) return nfft2d_fail(NFFT2D_ERR_ARG, "null data pointer");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    return p->dtype == NFFT2D_C64 ? t2_grid<float>(p, data, s) : t2_grid<double>(p, data, s);
}

extern "C" int nfft2d_type2_interp(nfft2d_plan *p, void *result, void *stream)
{
    int rc = check_exec(p, result, result);
    if (rc || p->M == 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    return p->dtype == NFFT2D_C64 ? gather<float, 0>(p, result, nullptr, nullptr, nullptr, nullptr, nullptr, s)
                                  : gather<double, 0>(p, result, nullptr, nullptr, nullptr, nullptr, nullptr, s);
}

extern "C" int nfft2d_type1_spread(nfft2d_plan *p, const void *data, void *stream)
{
    int rc = check_exec(p, data, data);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    return p->dtype == NFFT2D_C64 ? spread_impl<float>(p, data, false, s) : spread_impl<double>(p, data, false, s);
}

extern "C" int nfft2d_type1_finish(nfft2d_plan *p, void *result, int sign, void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "null plan");
    if (!result) return nfft2d_fail(NFFT2D_ERR_ARG, "null result pointer");
    if (sign != -1 && sign != 1) return nfft2d_fail(NFFT2D_ERR_ARG, "sign must be -1 or +1");
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (p->dtype == NFFT2D_C64)
        return sign < 0 ? t1_passes<float, -1>(p, result, s) : t1_passes<float, +1>(p, result, s);
    return sign < 0 ? t1_passes<double, -1>(p, result, s) : t1_passes<double, +1>(p, result, s);
}

extern "C" int nfft2d_type2(nfft2d_plan *p, void *result, const void *data, void *stream)
{
    int rc = check_exec(p, result, data);
    if (rc || p->M == 0) return rc;
    if ((rc = nfft2d_type2_grid(p, data, stream))) return rc;
    return nfft2d_type2_interp(p, result, stream);
}

// setpts + type-2 with the bin sort on the plan's second stream, concurrent
// with the grid passes on `stream`; the gather waits for both.
extern "C" int nfft2d_setpts_type2(nfft2d_plan *p, void *result, const void *data, const void *x, const void *y, int M,
                                   void *stream)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "setpts_type2: null plan");
    if (M < 0 || M > kMaxPoints)
        return nfft2d_fail(NFFT2D_ERR_ARG, "setpts_type2: M = %d outside [0, %d]", M, kMaxPoints);
    if (M > 0 && (!x || !y || !data || !result))
        return nfft2d_fail(NFFT2D_ERR_ARG, "setpts_type2: null data, result or coordinate pointer");
    if (M == 0) return nfft2d_setpts(p, x, y, 0, stream);
    int rc = nfft2d_check_device(p->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = p->fk.init(p->device);
    if (rc || (rc = p->fk.fork(s))) return rc;
    rc = nfft2d_setpts(p, x, y, M, p->fk.aux);
    const int rg = rc ? NFFT2D_OK : nfft2d_type2_grid(p, data, stream);
    const int rj = p->fk.join(s);               // always rejoin, so that a capture stays well formed
    if (rc || (rc = rg) || (rc = rj)) return rc;
    return nfft2d_type2_interp(p, result, stream);
}

extern "C" int nfft2d_type2_dx(nfft2d_plan *p, void *y, void *dy1, void *dy2, const void *data, void *stream)
{
    int rc = check_exec(p, y, data);
    if (rc || p->M == 0) return rc;
    if (!dy1 || !dy2) return nfft2d_fail(NFFT2D_ERR_ARG, "type2_dx: null derivative output");
    hipStream_t s = (hipStream_t)stream;
    if (p->dtype == NFFT2D_C64) {
        if ((rc = t2_grid<float>(p, data, s))) return rc;
        return gather<float, 1>(p, y, dy1, dy2, nullptr, nullptr, nullptr, s);
    }
    if ((rc = t2_grid<double>(p, data, s))) return rc;
    return gather<double, 1>(p, y, dy1, dy2, nullptr, nullptr, nullptr, s);
}

// Least squares in one gather: y = A(x) data, r = 2 (y - y_obs) and the
// coordinate gradient of |A(x) data - y_obs|^2, Re(conj(r) dy/dx) (the value
// and derivative sums share their tile reads).  r also goes, in the plan's
// sorted point order, into the plan's residual buffer for nfft2d_lsq_adjoint;
// the caller's r may be null.
extern "C" int nfft2d_type2_lsq(nfft2d_plan *p, void *r, void *gx1, void *gx2, const void *y_obs, const void *data,
                                void *stream)
{
    int rc = check_exec(p, data, data);
    if (rc || p->M == 0) return rc;
    if (!gx1 || !gx2 || !y_obs) return nfft2d_fail(NFFT2D_ERR_ARG, "type2_lsq: null argument");
    hipStream_t s = (hipStream_t)stream;
    if (p->M > p->residcap) {
        hfree(p->resid);
        p->resid = nullptr;
        p->residcap = 0;
        p->resid_valid = false;
        if (nfft2d_malloc(&p->resid, p->esz * (size_t)p->Mcap) != hipSuccess)
            return nfft2d_fail(NFFT2D_ERR_ALLOC, "type2_lsq: hipMalloc of the residual (%d points) failed", p->Mcap);
        p->residcap = p->Mcap;
    }
    if (p->dtype == NFFT2D_C64) {
        if ((rc = t2_grid<float>(p, data, s)) || (rc = gather<float, 3>(p, r, p->resid, nullptr, y_obs, gx1, gx2, s)))
            return rc;
    } else {
        if ((rc = t2_grid<double>(p, data, s)) || (rc = gather<double, 3>(p, r, p->resid, nullptr, y_obs, gx1, gx2, s)))
            return rc;
    }
    p->resid_valid = true;
    return NFFT2D_OK;
}

// The adjoint of the residual that the last nfft2d_type2_lsq left in the
// plan: dL/d(data) of the least-squares loss.  The spread reads the residual
// in the plan's sorted point order (no gather by original index); otherwise
// it is nfft2d_adjoint(r) (the same bits, except on grids under 44 x 44,
// whose spread sums with LDS atomics).
extern "C" int nfft2d_lsq_adjoint(nfft2d_plan *p, void *result, void *stream)
{
    int rc = check_exec(p, result, result);
    if (rc) return rc;
    if (!result) return nfft2d_fail(NFFT2D_ERR_ARG, "lsq_adjoint: null result pointer");
    if (p->M > 0 && !p->resid_valid)
        return nfft2d_fail(NFFT2D_ERR_STATE, "lsq_adjoint: no residual for the current points (call nfft2d_type2_lsq)");
    hipStream_t s = (hipStream_t)stream;
    if (p->dtype == NFFT2D_C64) {
        if ((rc = spread_impl<float>(p, p->resid, true, s))) return rc;
        return t1_passes<float, +1>(p, result, s);
    }
    if ((rc = spread_impl<double>(p, p->resid, true, s))) return rc;
    return t1_passes<double, +1>(p, result, s);
}

// The derivative gather of nfft2d_type2_grad_x on the oversampled grid the
// plan holds (left by nfft2d_type2 / nfft2d_type2_grid of the same data, with
// no spread since): a backward that runs right after its forward skips
// recomputing the grid.
extern "C" int nfft2d_type2_grad_x_grid(nfft2d_plan *p, void *gx1, void *gx2, const void *g, void *stream)
{
    int rc = check_exec(p, gx1, g);
    if (rc || p->M == 0) return rc;
    if (!gx2) return nfft2d_fail(NFFT2D_ERR_ARG, "type2_grad_x_grid: null argument");
    hipStream_t s = (hipStream_t)stream;
    return p->dtype == NFFT2D_C64 ? gather<float, 2>(p, nullptr, nullptr, nullptr, g, gx1, gx2, s)
                                  : gather<double, 2>(p, nullptr, nullptr, nullptr, g, gx1, gx2, s);
}

extern "C" int nfft2d_type2_grad_x(nfft2d_plan *p, void *gx1, void *gx2, const void *g, const void *data, void *stream)
{
    int rc = check_exec(p, gx1, data);
    if (rc || p->M == 0) return rc;
    if (!gx2 || !g) return nfft2d_fail(NFFT2D_ERR_ARG, "type2_grad_x: null argument");
    hipStream_t s = (hipStream_t)stream;
    if (p->dtype == NFFT2D_C64) {
        if ((rc = t2_grid<float>(p, data, s))) return rc;
        return gather<float, 2>(p, nullptr, nullptr, nullptr, g, gx1, gx2, s);
    }
    if ((rc = t2_grid<double>(p, data, s))) return rc;
    return gather<double, 2>(p, nullptr, nullptr, nullptr, g, gx1, gx2, s);
}

static int type1_common(nfft2d_plan *p, void *result, const void *data, int sign, void *stream)
{
    int rc = check_exec(p, data, data);
    if (rc) return rc;
    if (!result) return nfft2d_fail(NFFT2D_ERR_ARG, "null result pointer");
    if ((rc = nfft2d_type1_spread(p, data, stream))) return rc;
    return nfft2d_type1_finish(p, result, sign, stream);
}

extern "C" int nfft2d_type1(nfft2d_plan *p, void *result, const void *data, void *stream)
{
    return type1_common(p, result, data, -1, stream);
}

extern "C" int nfft2d_adjoint(nfft2d_plan *p, void *result, const void *data, void *stream)
{
    return type1_common(p, result, data, +1, stream);
}

// ============================================================== drop-in entry points
namespace {
std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, int>, nfft2d_plan *> g_cache;

int cached_plan(int N1, int N2, int dtype, nfft2d_plan **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "hipGetDevice failed");
    auto key = std::make_tuple(N1, N2, dtype, dev);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) { *out = it->second; return NFFT2D_OK; }
    int rc = nfft2d_plan_create(out, N1, N2, dtype, kK, kC, 0);
    if (rc) return rc;
    g_cache[key] = *out;
    return NFFT2D_OK;
}

int dropin(int dtype, bool type2, void *result, const void *data, const void *x, const void *y, int N1, int N2, int M)
{
    if (M < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "M < 0");
    if (!result) return nfft2d_fail(NFFT2D_ERR_ARG, "null result pointer");
    if (M > 0 && (!data || !x || !y)) return nfft2d_fail(NFFT2D_ERR_ARG, "null data or coordinate pointer");
    std::lock_guard<std::mutex> lk(g_cache_mu);
    nfft2d_plan *p = nullptr;
    int rc = cached_plan(N1, N2, dtype, &p);
    if (rc) return rc;
    if (type2) return nfft2d_setpts_type2(p, result, data, x, y, M, nullptr);
    if ((rc = nfft2d_setpts(p, x, y, M, nullptr))) return rc;
    return nfft2d_type1(p, result, data, nullptr);
}
}  // namespace

extern "C" int NFFT2_2D_GPU_c64(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y, int N1,
                                int N2, int M)
{
    return dropin(NFFT2D_C64, false, result, data, x, y, N1, N2, M);
}
extern "C" int NFFT2_2D_GPU_c128(nfft2d_c128 *result, const nfft2d_c128 *data, const double *x, const double *y,
                                 int N1, int N2, int M)
{
    return dropin(NFFT2D_C128, false, result, data, x, y, N1, N2, M);
}
extern "C" int NFFT1_2D_GPU_c64(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y, int N1,
                                int N2, int M)
{
    return dropin(NFFT2D_C64, true, result, data, x, y, N1, N2, M);
}
extern "C" int NFFT1_2D_GPU_c128(nfft2d_c128 *result, const nfft2d_c128 *data, const double *x, const double *y,
                                 int N1, int N2, int M)
{
    return dropin(NFFT2D_C128, true, result, data, x, y, N1, N2, M);
}

extern "C" void Calculate_cuFFT_plan_C2C_NFFT2_2D(int N1, int N2)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    nfft2d_plan *p;
    cached_plan(N1, N2, NFFT2D_C64, &p);
}
extern "C" void Calculate_cuFFT_plan_Z2Z_NFFT2_2D(int N1, int N2)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    nfft2d_plan *p;
    cached_plan(N1, N2, NFFT2D_C128, &p);
}
extern "C" void Destroy_cuFFT_plan_NFFT2_2D(void)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    (void)hipDeviceSynchronize();
    for (auto &kv : g_cache) free_plan(kv.second);
    g_cache.clear();
}

// ============================================================================
// NFFT2_2D_GPU_c64_noalloc: the complex64 type-1 drop-in for callers that
// cannot allocate device memory (code under graph capture, serving loops).
// Every buffer of the transform (tables, oversampled grid, sorted points,
// histograms, subproblem lists, spread tiles) is carved out of one
// file-scope device array, `scratch`, which each GPU holds from the moment
// the library is loaded on it until the process ends.  Its device address is
// taken once per device with hipGetSymbolAddress (the array's name is a host
// stand-in, not a device address) and kept in g_noalloc_ptr.  The kernels
// and their order are those of NFFT2_2D_GPU_c64 (the results agree to FP32
// rounding: the bin sort orders a bin's points as the waves reach its
// counter, which the spread's sums follow).  Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (different streams, or threads) race on it.  Every buffer is written
// before it is read within a call (the setpts passes write every histogram
// row, offset and subproblem entry; the spread writes every tile it merges and
// every grid cell), so nothing is cleared between calls.
constexpr int kNoAllocMaxN = 256;                   // N1, N2 <= 256
constexpr int kNoAllocMaxM = 1 << 18;               // M <= 262,144
constexpr int kNoAllocMaxL = kC * kNoAllocMaxN;
constexpr int kNoAllocMaxBins = (kNoAllocMaxL / kBin) * (kNoAllocMaxL / kBin);
constexpr int kNoAllocMaxSub = kNoAllocMaxBins + kNoAllocMaxM / kGatherThreads;
static_assert(kNoAllocMaxBins <= kMaxLdsBins, "the no-allocation path uses the LDS histogram (no rank array)");

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the buffers inside `scratch` (all at their maxima).
struct NoAllocLayout {
    size_t tw1, tw2, D1, D2, grid, tmp, pts, hist, tot, nsub, subbeg, gexcl, gsum, sub, tiles, total;
    constexpr NoAllocLayout() : tw1(0), tw2(0), D1(0), D2(0), grid(0), tmp(0), pts(0), hist(0), tot(0), nsub(0),
                                subbeg(0), gexcl(0), gsum(0), sub(0), tiles(0), total(0) {}
};
static NoAllocLayout noalloc_layout()
{
    NoAllocLayout o;
    size_t at = 0;
    auto take = [&](size_t &field, size_t bytes) { field = at; at += align256(bytes); };
    take(o.tw1, sizeof(float2) * 2 * kNoAllocMaxN);
    take(o.tw2, sizeof(float2) * 2 * kNoAllocMaxN);
    take(o.D1, sizeof(float) * kNoAllocMaxN);
    take(o.D2, sizeof(float) * kNoAllocMaxN);
    take(o.grid, sizeof(float2) * (size_t)kNoAllocMaxL * kNoAllocMaxL);
    take(o.tmp, sizeof(float2) * (size_t)kNoAllocMaxN * kNoAllocMaxL);
    take(o.pts, sizeof(PtRec<float>) * (size_t)kNoAllocMaxM);
    take(o.hist, sizeof(int) * (size_t)256 * kNoAllocMaxBins);      // at most 256 histogram blocks
    take(o.tot, sizeof(int) * kNoAllocMaxBins);
    take(o.nsub, sizeof(int));
    take(o.subbeg, sizeof(int) * (kNoAllocMaxBins + 1));
    take(o.gexcl, sizeof(int2) * kNoAllocMaxBins);
    take(o.gsum, sizeof(int2) * ((kNoAllocMaxBins + kColBins - 1) / kColBins));
    take(o.sub, sizeof(int4) * kNoAllocMaxSub);
    take(o.tiles, sizeof(float2) * kTileCells * (size_t)kNoAllocMaxSub);
    o.total = at;
    return o;
}
constexpr size_t kNoAllocBytes = (size_t)22 << 20;  // 22 MiB >= noalloc_layout().total (checked at set-up)

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};       // device address of `scratch` per device

// Host copies of the complex64 twiddle / deconvolution tables per N (the
// values upload_tables<float> computes), kept so that an asynchronous copy
// never reads freed host memory.
std::map<int, std::pair<std::vector<float2>, std::vector<float>>> g_noalloc_tables;
WinArgs g_noalloc_wa;
bool g_noalloc_wa_ready = false;

const std::pair<std::vector<float2>, std::vector<float>> &noalloc_tables(int N)
{
    auto it = g_noalloc_tables.find(N);
    if (it != g_noalloc_tables.end()) return it->second;
    const double S = host_i0(kK * host_alpha()), a = host_alpha();
    std::vector<float2> tw(2 * N);
    for (int q = 0; q < 2 * N; ++q) {
        const double ang = -kPi * (double)q / (double)N;
        tw[q] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    std::vector<float> D(N);
    for (int i = 0; i < N; ++i) {
        const double xi = 2.0 * kPi * (double)(i - N / 2) / ((double)kC * N);
        D[i] = (float)(S / host_i0(kK * std::sqrt(a * a - xi * xi)));
    }
    return g_noalloc_tables[N] = {std::move(tw), std::move(D)};
}

// The device address of `scratch` on the current device, looked up once per device.
int noalloc_base(int device, unsigned char **out)
{
    if (device < 0 || device >= kNoAllocMaxDevices)
        return nfft2d_fail(NFFT2D_ERR_ARG, "NFFT2_2D_GPU_c64_noalloc: device %d above the supported %d", device,
                           kNoAllocMaxDevices);
    if (!g_noalloc_ptr[device]) {
        if (noalloc_layout().total > kNoAllocBytes)
            return nfft2d_fail(NFFT2D_ERR_SIZE, "NFFT2_2D_GPU_c64_noalloc: scratch layout exceeds the array");
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return nfft2d_fail(NFFT2D_ERR_HIP, "NFFT2_2D_GPU_c64_noalloc: hipGetSymbolAddress(scratch): %s",
                               hipGetErrorString(e));
        g_noalloc_ptr[device] = ptr;
    }
    *out = static_cast<unsigned char *>(g_noalloc_ptr[device]);
    return NFFT2D_OK;
}
}  // namespace

extern "C" int NFFT2_2D_GPU_c64_noalloc(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y,
                                        int N1, int N2, int M, void *stream)
{
    if (M < 0) return nfft2d_fail(NFFT2D_ERR_ARG, "NFFT2_2D_GPU_c64_noalloc: M < 0");
    if (!result) return nfft2d_fail(NFFT2D_ERR_ARG, "NFFT2_2D_GPU_c64_noalloc: null result pointer");
    if (M > 0 && (!data || !x || !y)) return nfft2d_fail(NFFT2D_ERR_ARG, "NFFT2_2D_GPU_c64_noalloc: null pointer");
    if (N1 <= 0 || N2 <= 0 || (N1 & 1) || (N2 & 1) || N1 > kNoAllocMaxN || N2 > kNoAllocMaxN)
        return nfft2d_fail(NFFT2D_ERR_SIZE,
                           "NFFT2_2D_GPU_c64_noalloc: N1, N2 must be even and at most %d (got %d, %d); larger "
                           "transforms go through NFFT2_2D_GPU_c64",
                           kNoAllocMaxN, N1, N2);
    if (M > kNoAllocMaxM)
        return nfft2d_fail(NFFT2D_ERR_SIZE,
                           "NFFT2_2D_GPU_c64_noalloc: M = %d above the maximum %d; larger point sets go through "
                           "NFFT2_2D_GPU_c64",
                           M, kNoAllocMaxM);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char *base = nullptr;
    int rc = noalloc_base(device, &base);
    if (rc) return rc;
    const NoAllocLayout o = noalloc_layout();
    // a plan whose buffers are views of `scratch`, sized to the maxima so that
    // no path of setpts / spread grows (allocates) anything
    nfft2d_plan_s p;
    p.N1 = N1; p.N2 = N2; p.L1 = kC * N1; p.L2 = kC * N2; p.dtype = NFFT2D_C64; p.device = device;
    p.esz = sizeof(float2);
    p.nb1 = (p.L1 + kBin - 1) / kBin;
    p.nb2 = (p.L2 + kBin - 1) / kBin;
    p.nbins = p.nb1 * p.nb2;
    p.geo.L1 = p.L1; p.geo.L2 = p.L2; p.geo.nb1 = p.nb1;
    p.geo.inv1d = 1.0 / p.L1; p.geo.inv2d = 1.0 / p.L2;
    p.geo.inv1f = (float)p.geo.inv1d; p.geo.inv2f = (float)p.geo.inv2d;
    if (!g_noalloc_wa_ready) {
        make_window_args(g_noalloc_wa);
        g_noalloc_wa_ready = true;
    }
    p.wa = g_noalloc_wa;
    p.tw1 = base + o.tw1; p.tw2 = base + o.tw2; p.D1 = base + o.D1; p.D2 = base + o.D2;
    p.grid = base + o.grid; p.tmp = base + o.tmp; p.pts = base + o.pts;
    p.hist = reinterpret_cast<int *>(base + o.hist); p.tot = reinterpret_cast<int *>(base + o.tot);
    p.nsub = reinterpret_cast<int *>(base + o.nsub); p.subbeg = reinterpret_cast<int *>(base + o.subbeg);
    p.gexcl = reinterpret_cast<int2 *>(base + o.gexcl); p.gsum = reinterpret_cast<int2 *>(base + o.gsum);
    p.sub = reinterpret_cast<int4 *>(base + o.sub); p.scratch = base + o.tiles;
    p.Mcap = kNoAllocMaxM; p.subcap = kNoAllocMaxSub; p.histcap = 256 * kNoAllocMaxBins; p.scratchcap = kNoAllocMaxSub;
    p.rank = nullptr;                                  // the LDS histogram path needs none
    hipStream_t s = (hipStream_t)stream;
    for (int dim = 0; dim < 2; ++dim) {
        const int N = dim == 0 ? N1 : N2;
        const auto &tb = noalloc_tables(N);
        NFFT2D_HIP_TRY(hipMemcpyAsync(dim == 0 ? p.tw1 : p.tw2, tb.first.data(), sizeof(float2) * 2 * N,
                                      hipMemcpyHostToDevice, s));
        NFFT2D_HIP_TRY(hipMemcpyAsync(dim == 0 ? p.D1 : p.D2, tb.second.data(), sizeof(float) * N,
                                      hipMemcpyHostToDevice, s));
    }
    // the same calls as NFFT2_2D_GPU_c64: setpts, spread, FFT + crop + deconvolution
    if ((rc = setpts_impl<float>(&p, x, y, M, s))) return rc;
    if ((rc = spread_impl<float>(&p, data, false, s))) return rc;
    return t1_passes<float, -1>(&p, result, s);
}

// ============================================================================
// Batched 1-D NFFTs (complex64): BP_FBP_FFBP/NFFT1_1D.m (type-2) and
// NFFT/NFFT2/1D/Matlab/NFFT2_1D.m (type-1), B lines per call (kernels in
// nfft1d_kernels.hpp; the row passes are the 2-D transform's).
struct nfft1d_plan_s {
    int N = 0, L = 0, device = 0, rows = 0;   // rows: lines the grid and the unit row factors hold
    void *tw = nullptr, *D = nullptr;         // 2N twiddles, N deconvolution factors S / phi(xi_k)
    void *ones = nullptr, *grid = nullptr;    // rows x 1.0f (the row passes' per-row factor), rows x L grid
    void *czt = nullptr;                      // chirp-z tables when N is not a power of two (from 128 on)
    int cztm = 0;
    WinArgs wa;
};

static void free_plan1d(nfft1d_plan *p)
{
    if (!p) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(p->device);
    hfree(p->tw); hfree(p->D); hfree(p->ones); hfree(p->grid); hfree(p->czt);
    (void)hipSetDevice(cur);
    delete p;
}

extern "C" int nfft1d_plan_create(nfft1d_plan **out, int N)
{
    if (!out) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft1d_plan_create: null output pointer");
    *out = nullptr;
    if (N < 2 || (N & 1) || N > 4096 || (!is_pow2(N) && N > 2048))
        return nfft2d_fail(NFFT2D_ERR_SIZE,
                           "nfft1d_plan_create: N must be even, at most 2048, or a power of two up to 4096 (got %d)", N);
    int device = 0;
    NFFT2D_HIP_TRY(hipGetDevice(&device));
    nfft1d_plan *p = new (std::nothrow) nfft1d_plan_s();
    if (!p) return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft1d_plan_create: host allocation failed");
    p->N = N;
    p->L = kC * N;
    p->device = device;
    make_window_args(p->wa);
    const auto &tb = noalloc_tables(N);        // the complex64 tables of upload_tables<float>
    if (nfft2d_malloc(&p->tw, sizeof(float2) * 2 * N) != hipSuccess || nfft2d_malloc(&p->D, sizeof(float) * N) != hipSuccess) {
        free_plan1d(p);
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft1d_plan_create: hipMalloc of the tables failed");
    }
    if (hipMemcpy(p->tw, tb.first.data(), sizeof(float2) * 2 * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->D, tb.second.data(), sizeof(float) * N, hipMemcpyHostToDevice) != hipSuccess) {
        free_plan1d(p);
        return nfft2d_fail(NFFT2D_ERR_HIP, "nfft1d_plan_create: upload of the tables failed");
    }
    if (czt_path(N, NFFT2D_C64)) {
        const int rc = czt_buffer<float>(N, &p->czt, &p->cztm);
        if (rc) {
            free_plan1d(p);
            return rc;
        }
    }
    *out = p;
    return NFFT2D_OK;
}

extern "C" int nfft1d_plan_destroy(nfft1d_plan *p)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "nfft1d_plan_destroy: null plan");
    free_plan1d(p);
    return NFFT2D_OK;
}

// Grid and unit row factors for B lines (grown, never shrunk; the first call
// with more lines than before allocates).
static int plan1d_rows(nfft1d_plan *p, int B)
{
    if (B <= p->rows) return NFFT2D_OK;
    hfree(p->ones); hfree(p->grid);
    p->ones = p->grid = nullptr;
    p->rows = 0;
    if (nfft2d_malloc(&p->ones, sizeof(float) * B) != hipSuccess ||
        nfft2d_malloc(&p->grid, sizeof(float2) * (size_t)B * p->L) != hipSuccess)
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "nfft1d: hipMalloc for %d lines failed", B);
    std::vector<float> one(B, 1.0f);
    NFFT2D_HIP_TRY(hipMemcpy(p->ones, one.data(), sizeof(float) * B, hipMemcpyHostToDevice));
    p->rows = B;
    return NFFT2D_OK;
}

static int check1d(nfft1d_plan *p, const void *a, const void *b, const void *x, int M, int B, bool type2,
                   const char *what)
{
    if (!p) return nfft2d_fail(NFFT2D_ERR_ARG, "%s: null plan", what);
    if (M < 0 || B < 0 || M > kMaxPoints) return nfft2d_fail(NFFT2D_ERR_ARG, "%s: bad M = %d or batch = %d", what, M, B);
    if ((long long)M * B > (1LL << 31) - 1)
        return nfft2d_fail(NFFT2D_ERR_ARG, "%s: batch x M above 2^31 - 1", what);
    // a: the output (type-2: M per line, type-1: N), b: the input (type-2: N, type-1: M per line)
    const bool need_a = B > 0 && (!type2 || M > 0), need_b = B > 0 && (type2 || M > 0);
    if ((need_a && !a) || (need_b && !b) || (B > 0 && M > 0 && !x))
        return nfft2d_fail(NFFT2D_ERR_ARG, "%s: null pointer", what);
    return nfft2d_check_device(p->device);
}

extern "C" int nfft1d_type2(nfft1d_plan *p, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, int M,
                            int batch, void *stream)
{
    int rc = check1d(p, result, data, x, M, batch, true, "nfft1d_type2");
    if (rc || batch == 0 || M == 0) return rc;
    if ((rc = plan1d_rows(p, batch))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const float2 *f = reinterpret_cast<const float2 *>(data);
    float2 *G = (float2 *)p->grid;
    const float *ones = (const float *)p->ones, *D = (const float *)p->D;
    const float2 *tw = (const float2 *)p->tw;
    if (fast_path(p->N, NFFT2D_C64)) {
        dispatch_n(p->N, [&](auto n) {
            constexpr int N = decltype(n)::value;
            constexpr int R = rows_per_block<N>();
            hipLaunchKernelGGL((k_t2_rows<float, N, -1>), dim3(grid1d(batch, R)), dim3(R * (N / 8)), 0, s, f, G, ones, D,
                               tw, batch);
        });
    } else if (p->czt) {
        if ((rc = czt_launch<float, -1, true>(czt_tables_of<float>(p->czt, p->N, p->cztm, true), batch, f, p->N, 1, G,
                                              p->L, 1, ones, D, s)))
            return rc;
    } else {
        hipLaunchKernelGGL((k_t2_rows_direct<float, -1>), dim3(batch), dim3(256), sizeof(float2) * 3 * p->N, s, f, G, ones, D,
                           tw, batch, p->N);
    }
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gather1d, dim3(batch, grid1d(M, k1dPointsPerBlock)), dim3(k1dThreads), sizeof(float2) * p->L, s,
                       G, p->L, 1.0f / p->L, x, M, p->wa, reinterpret_cast<float2 *>(result));
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft1d_type1(nfft1d_plan *p, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, int M,
                            int batch, void *stream)
{
    int rc = check1d(p, result, data, x, M, batch, false, "nfft1d_type1");
    if (rc || batch == 0) return rc;
    if (p->N > 2048) return nfft2d_fail(NFFT2D_ERR_SIZE, "nfft1d_type1: N above 2048 (its LDS accumulators)");
    if ((rc = plan1d_rows(p, batch))) return rc;
    hipStream_t s = (hipStream_t)stream;
    float2 *G = (float2 *)p->grid, *F = reinterpret_cast<float2 *>(result);
    const float *ones = (const float *)p->ones, *D = (const float *)p->D;
    const float2 *tw = (const float2 *)p->tw;
    hipLaunchKernelGGL(k_spread1d, dim3(batch), dim3(k1dThreads), sizeof(unsigned long long) * 2 * p->L, s,
                       reinterpret_cast<const float2 *>(data), x, M, p->L, 1.0f / p->L, p->wa, G);
    NFFT2D_LAUNCH_CHECK();
    if (fast_path(p->N, NFFT2D_C64)) {
        dispatch_n(p->N, [&](auto n) {
            constexpr int N = decltype(n)::value;
            constexpr int R = rows_per_block<N>();
            hipLaunchKernelGGL((k_t1_rows<float, N, -1>), dim3(grid1d(batch, R)), dim3(R * (N / 8)), 0, s, G, F, ones, D,
                               tw, batch);
        });
    } else if (p->czt) {
        if ((rc = czt_launch<float, -1, false>(czt_tables_of<float>(p->czt, p->N, p->cztm, false), batch, G, p->L, 1,
                                               F, p->N, 1, ones, D, s)))
            return rc;
    } else {
        hipLaunchKernelGGL((k_t1_rows_direct<float, -1>), dim3(batch), dim3(256), sizeof(float2) * 2 * p->L, s, G, F, ones,
                           D, tw, batch, p->N);
    }
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// ============================================================================
// Batches: many complex64 type-2 NFFTs of different sizes, one launch per stage
// (kernels in nfft2d_batch.hpp).
#include "nfft2d_batch.hpp"

static constexpr int kBatchGatherWaveBlocks = 256 * 4;   // 4 blocks of 4 waves per CU (LDS: 38 KB each)

struct nfft2d_batch_s {
    int count = 0, device = 0;
    std::vector<nfft2d_plan *> plans;
    std::vector<BatchDesc> hdesc;
    BatchDesc *ddesc = nullptr;
    int2 *own = nullptr;                  // owner tables of the 5 per-transform stages, back to back
    int off_hist = 0, n_hist = 0, off_col = 0, n_col = 0;   // the scatter takes the histogram's blocks
    int off_rows = 0, n_rows = 0, off_cols = 0, n_cols = 0, off_merge = 0, n_merge = 0;
    // type-1 passes: the transforms whose lines go to the matrix cores (t1_pass_on_mfma)
    // in their own stages, the others in stages of the type-2 passes' shape
    int off_t1c = 0, n_t1c = 0, off_t1cm = 0, n_t1cm = 0, off_t1r = 0, n_t1r = 0, off_t1rm = 0, n_t1rm = 0;
    int *base = nullptr;                  // count + 1: subproblem offsets, base[count] = total
    int4 *bsub = nullptr;                 // the batch's subproblem list
    long long bsubcap = 0;
    int max_nbins = 0;
    bool have_points = false, have_io = false, have_io1 = false;
    bool sorted = false;                  // the subproblem list matches the points and sizes (batch_setpts ran)
    Fork fk;                              // setpts beside the grid passes (nfft2d_batch_setpts_type2)
    std::vector<int> dmax;                // per transform: direct-gather cut-off (BatchDesc::dmax)
    int gather_blocks = kBatchGatherWaveBlocks;   // width of the persistent wave gather
};

static int batch_rows_per_block(int N) { return N >= 1024 ? 1 : 1024 / N; }
static int batch_cols_per_block(int N) { return N >= 2048 ? 1 : 2048 / N; }
static constexpr int kBatchMinMaxpts = 64;   // smallest subproblem size the autotuner tries

static void batch_free(nfft2d_batch *b)
{
    if (!b) return;
    for (auto *p : b->plans) free_plan(p);   // null for plans a failed create never made
    hfree(b->ddesc); hfree(b->own); hfree(b->base); hfree(b->bsub);
    b->fk.destroy();
    delete b;
}

// Histogram blocking of a batch's transform: chunks of kBatchHistChunk points
// (the plans' 2048-point chunks fill the chip for one transform; a batch's
// transforms already do, and each (block, bin) count is 4 bytes written,
// scanned and read back: a 4096-bin row per 2048 points made the count matrix
// as large as the coordinates).  At most the plan's block count, so the plan's
// count matrix holds it.
#ifndef NFFT2D_BATCH_HIST_CHUNK
#define NFFT2D_BATCH_HIST_CHUNK 8192
#endif
static constexpr int kBatchHistChunk = NFFT2D_BATCH_HIST_CHUNK;
static void batch_blocking(const nfft2d_plan *p, int M, int &nblk, int &chunk)
{
    setpts_blocking(p, M, nblk, chunk);
    if (p->nbins <= kMaxLdsBins) {
        nblk = std::min(nblk, std::max(1, (M + kBatchHistChunk - 1) / kBatchHistChunk));
        chunk = std::max(1, (M + nblk - 1) / nblk);
    }
}

// desc[t] from the plan (sizes, buffers); the caller's pointers are kept.
static void batch_fill_desc(nfft2d_batch *b, int t)
{
    nfft2d_plan *p = b->plans[t];
    BatchDesc &d = b->hdesc[t];
    d.N1 = p->N1; d.N2 = p->N2; d.L1 = p->L1; d.L2 = p->L2; d.nb1 = p->nb1; d.nbins = p->nbins;
    d.maxpts = p->maxpts;
    batch_blocking(p, d.M, d.nblk, d.chunk);
    d.ngroups = (int)grid1d(p->nbins, kColBins);
    d.inv1 = (float)(1.0 / p->L1); d.inv2 = (float)(1.0 / p->L2);
    d.geo = p->geo;
    d.Tm = (float2 *)p->tmp; d.Gt = (float2 *)p->grid;
    d.D1 = (const float *)p->D1; d.D2 = (const float *)p->D2;
    d.tw1 = (const float2 *)p->tw1; d.tw2 = (const float2 *)p->tw2;
    d.hist = p->hist; d.rank = p->rank; d.tot = p->tot; d.subbeg = p->subbeg; d.nsub = p->nsub;
    d.gexcl = p->gexcl; d.gsum = p->gsum; d.sub = p->sub; d.pts = (PtRec<float> *)p->pts;
    d.scratch = (float2 *)p->scratch;
    d.r1 = merge_radius(p->L1);
    d.r2 = merge_radius(p->L2);
    d.dmax = b->dmax.empty() ? kDirectDefault : b->dmax[t];
    // the plan's chirp-z tables when their FFT fits the batch passes' LDS (longer
    // lines take the direct passes in a batch)
    d.czt1 = p->czt[0] && p->cztm[0] <= kBatchCztMaxM ? p->czt[0] : nullptr;
    d.czt2 = p->czt[1] && p->cztm[1] <= kBatchCztMaxM ? p->czt[1] : nullptr;
    d.cztm1 = p->cztm[0];
    d.cztm2 = p->cztm[1];
}

static int batch_upload(nfft2d_batch *b)
{
    NFFT2D_HIP_TRY(hipMemcpy(b->ddesc, b->hdesc.data(), sizeof(BatchDesc) * b->count, hipMemcpyHostToDevice));
    return NFFT2D_OK;
}

extern "C" int nfft2d_batch_create(nfft2d_batch **out, int count, const int *N1, const int *N2, const int *M,
                                   unsigned flags)
{
    if (!out) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_create: null batch pointer");
    *out = nullptr;
    if (count <= 0 || !N1 || !N2 || !M) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_create: count <= 0 or null sizes");
    if (flags != 0) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_create: flags must be 0");
    for (int t = 0; t < count; ++t) {
        if (N1[t] < 2 || N2[t] < 2 || (N1[t] & 1) || (N2[t] & 1) || N1[t] > kBatchMaxN || N2[t] > kBatchMaxN)
            return nfft2d_fail(NFFT2D_ERR_SIZE, "batch_create: transform %d: N1=%d N2=%d must be even and in [2, %d]",
                               t, N1[t], N2[t], kBatchMaxN);
        if (M[t] < 0 || M[t] > kMaxPoints)
            return nfft2d_fail(NFFT2D_ERR_ARG, "batch_create: transform %d: M = %d outside [0, %d]", t, M[t], kMaxPoints);
    }
    nfft2d_batch *b = new (std::nothrow) nfft2d_batch_s();
    if (!b) return nfft2d_fail(NFFT2D_ERR_ALLOC, "batch_create: host allocation failed");
    b->count = count;
    (void)hipGetDevice(&b->device);
    b->plans.assign(count, nullptr);
    b->hdesc.assign(count, BatchDesc{});
    b->dmax.assign(count, kDirectDefault);
    std::vector<int2> own;
    long long subcap = 0;
    for (int t = 0; t < count; ++t) {
        int rc = nfft2d_plan_create(&b->plans[t], N1[t], N2[t], NFFT2D_C64, 6, 2, 0);
        if (rc) { batch_free(b); return rc; }
        nfft2d_plan *p = b->plans[t];
        // buffers for M points at the smallest subproblem size the autotuner may pick
        p->maxpts = kBatchMinMaxpts;
        rc = ensure_point_buffers<float>(p, M[t]);
        p->maxpts = kGatherThreads;
        if (rc) { batch_free(b); return rc; }
        b->hdesc[t].M = M[t];
        batch_fill_desc(b, t);
        b->max_nbins = std::max(b->max_nbins, p->nbins);
        subcap += p->nbins + (M[t] + kBatchMinMaxpts - 1) / kBatchMinMaxpts;
    }
    auto add_stage = [&](int &off, int &n, auto blocks_of) {
        off = (int)own.size();
        for (int t = 0; t < b->count; ++t) {
            const int nb = blocks_of(t);
            for (int k = 0; k < nb; ++k) own.push_back(make_int2(t, k));
        }
        n = (int)own.size() - off;
    };
    add_stage(b->off_hist, b->n_hist, [&](int t) { return b->hdesc[t].nblk; });
    add_stage(b->off_col, b->n_col, [&](int t) { return b->hdesc[t].ngroups; });
    // (the scatter takes the histogram stage's blocks)
    // the Stockham passes take R rows / B columns per block, the direct ones one
    add_stage(b->off_rows, b->n_rows, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return batch_fft_n(d.N2) ? (int)grid1d(d.N1, batch_rows_per_block(d.N2)) : d.N1;
    });
    add_stage(b->off_cols, b->n_cols, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return batch_fft_n(d.N1) ? (int)grid1d(d.L2, batch_cols_per_block(d.N1)) : d.L2;
    });
    // type-1: 16-column tiles (column pass) and 16-row tiles (row pass) on the matrix cores
    add_stage(b->off_t1c, b->n_t1c, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return t1_pass_on_mfma(d.N1) ? 0 : batch_fft_n(d.N1) ? (int)grid1d(d.L2, batch_cols_per_block(d.N1)) : d.L2;
    });
    add_stage(b->off_t1cm, b->n_t1cm, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return t1_pass_on_mfma(d.N1) ? (int)grid1d(d.L2, 16) : 0;
    });
    add_stage(b->off_t1r, b->n_t1r, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return t1_pass_on_mfma(d.N2) ? 0 : batch_fft_n(d.N2) ? (int)grid1d(d.N1, batch_rows_per_block(d.N2)) : d.N1;
    });
    add_stage(b->off_t1rm, b->n_t1rm, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return t1_pass_on_mfma(d.N2) ? (int)grid1d(d.N1, 16) : 0;
    });
    // the tile merge, except where the matrix-core column pass sums the tiles itself
    add_stage(b->off_merge, b->n_merge, [&](int t) {
        const BatchDesc &d = b->hdesc[t];
        return t1_cols_merge_tiles(d.N1, d.r1, d.r2) ? 0 : d.nbins;
    });
    b->bsubcap = std::max<long long>(subcap, 1);
    if (nfft2d_malloc(&b->ddesc, sizeof(BatchDesc) * count) != hipSuccess ||
        nfft2d_malloc(&b->own, sizeof(int2) * own.size()) != hipSuccess ||
        nfft2d_malloc(&b->base, sizeof(int) * (count + 1)) != hipSuccess ||
        nfft2d_malloc(&b->bsub, sizeof(int4) * b->bsubcap) != hipSuccess) {
        (void)hipGetLastError();
        batch_free(b);
        return nfft2d_fail(NFFT2D_ERR_ALLOC, "batch_create: device allocation failed");
    }
    if (hipMemcpy(b->own, own.data(), sizeof(int2) * own.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->base, 0, sizeof(int) * (count + 1)) != hipSuccess || batch_upload(b) != NFFT2D_OK) {
        batch_free(b);
        return nfft2d_fail(NFFT2D_ERR_HIP, "batch_create: upload of the batch tables failed");
    }
    *out = b;
    return NFFT2D_OK;
}

extern "C" int nfft2d_batch_destroy(nfft2d_batch *b)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_destroy: null batch");
    batch_free(b);
    return NFFT2D_OK;
}

extern "C" int nfft2d_batch_count(const nfft2d_batch *b) { return b ? b->count : nfft2d_fail(NFFT2D_ERR_ARG, "null batch"); }

extern "C" int nfft2d_batch_set_points(nfft2d_batch *b, const float *const *x, const float *const *y)
{
    if (!b || !x || !y) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_points: null argument");
    for (int t = 0; t < b->count; ++t) {
        if (b->hdesc[t].M > 0 && (!x[t] || !y[t]))
            return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_points: transform %d: null coordinates", t);
        b->hdesc[t].x = x[t];
        b->hdesc[t].y = y[t];
    }
    b->have_points = true;
    b->sorted = false;
    return batch_upload(b);
}

extern "C" int nfft2d_batch_set_io(nfft2d_batch *b, const nfft2d_c64 *const *modes, nfft2d_c64 *const *result)
{
    if (!b || !modes || !result) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_io: null argument");
    for (int t = 0; t < b->count; ++t) {
        if (!modes[t] || (b->hdesc[t].M > 0 && !result[t]))
            return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_io: transform %d: null array", t);
        b->hdesc[t].f = reinterpret_cast<const float2 *>(modes[t]);
        b->hdesc[t].out = reinterpret_cast<float2 *>(result[t]);
    }
    b->have_io = true;
    return batch_upload(b);
}

extern "C" int nfft2d_batch_set_maxpts(nfft2d_batch *b, int t, int maxpts)
{
    if (!b || t < 0 || t >= b->count) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_maxpts: bad batch or index");
    if (maxpts < kBatchMinMaxpts || maxpts > kGatherThreads)
        return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_maxpts: maxpts must be in [%d, %d]", kBatchMinMaxpts,
                           kGatherThreads);
    b->plans[t]->maxpts = maxpts;
    b->hdesc[t].maxpts = maxpts;
    b->sorted = false;                    // the subproblem list must be rebuilt
    return batch_upload(b);
}

extern "C" int nfft2d_batch_get_maxpts(const nfft2d_batch *b, int t)
{
    if (!b || t < 0 || t >= b->count) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_get_maxpts: bad batch or index");
    return b->hdesc[t].maxpts;
}

extern "C" int nfft2d_batch_set_direct(nfft2d_batch *b, int t, int npts)
{
    if (!b || t < 0 || t >= b->count) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_direct: bad batch or index");
    if (npts < 0 || npts > kDirectCap)
        return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_direct: the cut-off must be in [0, %d]", kDirectCap);
    b->dmax[t] = npts;
    b->hdesc[t].dmax = npts;
    b->sorted = false;                    // the subproblem list carries the cut-off: rebuilt by the next setpts
    return batch_upload(b);
}

extern "C" int nfft2d_batch_get_direct(const nfft2d_batch *b, int t)
{
    if (!b || t < 0 || t >= b->count) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_get_direct: bad batch or index");
    return b->dmax[t];
}

extern "C" int nfft2d_batch_set_gather_blocks(nfft2d_batch *b, int blocks)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_gather_blocks: null batch");
    if (blocks < 1 || blocks > 65536)
        return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_gather_blocks: %d outside [1, 65536]", blocks);
    b->gather_blocks = blocks;
    return NFFT2D_OK;
}

extern "C" int nfft2d_batch_get_gather_blocks(const nfft2d_batch *b)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_get_gather_blocks: null batch");
    return b->gather_blocks;
}

extern "C" int nfft2d_batch_setpts(nfft2d_batch *b, void *stream)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_setpts: null batch");
    if (!b->have_points) return nfft2d_fail(NFFT2D_ERR_STATE, "batch_setpts: nfft2d_batch_set_points was not called");
    int rc = nfft2d_check_device(b->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (b->n_hist > 0)
        hipLaunchKernelGGL(kb_hist, dim3(b->n_hist), dim3(kHistThreads), sizeof(int) * b->max_nbins, s, b->ddesc,
                           b->own + b->off_hist);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_colscan, dim3(b->n_col), dim3(256), 0, s, b->ddesc, b->own + b->off_col);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_scan_groups, dim3(b->count), dim3(kScanThreads), 0, s, b->ddesc);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_sub_prefix, dim3(1), dim3(1024), 0, s, b->ddesc, b->count, b->base);
    NFFT2D_LAUNCH_CHECK();
    // one scatter block per histogram block (the same owner table)
    if (b->n_hist > 0)
        hipLaunchKernelGGL(kb_scatter_lds, dim3(b->n_hist), dim3(kScatThreads), sizeof(int) * b->max_nbins, s,
                           b->ddesc, b->own + b->off_hist, b->base, b->bsub);
    NFFT2D_LAUNCH_CHECK();
    b->sorted = true;
    return NFFT2D_OK;
}

static int batch_grid(nfft2d_batch *b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_t2_rows, dim3(b->n_rows), dim3(128), 0, s, b->ddesc, b->own + b->off_rows);
    NFFT2D_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_t2_cols, dim3(b->n_cols), dim3(256), 0, s, b->ddesc, b->own + b->off_cols);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// The two gathers take disjoint subproblems and write disjoint outputs, but
// they run one after the other: the direct one on the second stream beside
// the wave gather made the batch step slower (2.25 against 1.83 ms).
static int batch_gather(nfft2d_batch *b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_gather_w, dim3(b->gather_blocks), dim3(256), 0, s, b->ddesc, b->bsub, b->base + b->count,
                       b->plans[0]->wa);
    NFFT2D_LAUNCH_CHECK();
    constexpr int per_block = 256 / kDirectLanes;   // list entries per block
    hipLaunchKernelGGL(kb_gather_direct, dim3((unsigned)((b->bsubcap + per_block - 1) / per_block)), dim3(256), 0, s,
                       b->ddesc, b->bsub, b->base + b->count, b->plans[0]->wa);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

extern "C" int nfft2d_batch_type2(nfft2d_batch *b, void *stream)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_type2: null batch");
    if (!b->have_io) return nfft2d_fail(NFFT2D_ERR_STATE, "batch_type2: nfft2d_batch_set_io was not called");
    if (!b->sorted)
        return nfft2d_fail(NFFT2D_ERR_STATE, "batch_type2: the points are not sorted (call nfft2d_batch_setpts after "
                                             "nfft2d_batch_set_points / nfft2d_batch_set_maxpts)");
    int rc = nfft2d_check_device(b->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = batch_grid(b, s);
    return rc ? rc : batch_gather(b, s);
}

// Type-1 of every transform: its M[t] samples data[t] at its sorted points
// to the N1 x N2 modes result[t] (NFFT2_2D_GPU per transform).  set_io_type1
// allocates each plan's spread tiles (the tile merge needs at least 3 bins
// and 44 cells per dimension: N1, N2 >= 22).
extern "C" int nfft2d_batch_set_io_type1(nfft2d_batch *b, const nfft2d_c64 *const *data, nfft2d_c64 *const *result)
{
    if (!b || !data || !result) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_io_type1: null argument");
    int rc = nfft2d_check_device(b->device);
    if (rc) return rc;
    for (int t = 0; t < b->count; ++t) {
        if (!result[t] || (b->hdesc[t].M > 0 && !data[t]))
            return nfft2d_fail(NFFT2D_ERR_ARG, "batch_set_io_type1: transform %d: null array", t);
        if (!mfma_spread_ok(b->plans[t]))
            return nfft2d_fail(NFFT2D_ERR_SIZE, "batch_set_io_type1: transform %d: N1, N2 must be at least 22", t);
    }
    for (int t = 0; t < b->count; ++t) {
        nfft2d_plan *p = b->plans[t];
        if (p->subcap > p->scratchcap) {
            free_scratch(p);
            if (nfft2d_malloc(&p->scratch, sizeof(float2) * kTileCells * (size_t)p->subcap) != hipSuccess)
                return nfft2d_fail(NFFT2D_ERR_ALLOC, "batch_set_io_type1: hipMalloc of the spread tiles failed");
            p->scratchcap = p->subcap;
        }
        b->hdesc[t].scratch = (float2 *)p->scratch;
        b->hdesc[t].d1 = reinterpret_cast<const float2 *>(data[t]);
        b->hdesc[t].F1 = reinterpret_cast<float2 *>(result[t]);
    }
    b->have_io1 = true;
    return batch_upload(b);
}

extern "C" int nfft2d_batch_type1(nfft2d_batch *b, void *stream)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_type1: null batch");
    if (!b->have_io1) return nfft2d_fail(NFFT2D_ERR_STATE, "batch_type1: nfft2d_batch_set_io_type1 was not called");
    if (!b->sorted)
        return nfft2d_fail(NFFT2D_ERR_STATE, "batch_type1: the points are not sorted (call nfft2d_batch_setpts after "
                                             "nfft2d_batch_set_points / nfft2d_batch_set_maxpts)");
    int rc = nfft2d_check_device(b->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(kb_spread, dim3((unsigned)b->bsubcap), dim3(256), 0, s, b->ddesc, b->bsub, b->base, b->count,
                       b->plans[0]->wa);
    NFFT2D_LAUNCH_CHECK();
    if (b->n_merge > 0)
        hipLaunchKernelGGL(kb_spread_merge, dim3(b->n_merge), dim3(256), 0, s, b->ddesc, b->own + b->off_merge);
    NFFT2D_LAUNCH_CHECK();
    // column passes (FFT / chirp-z / direct lines, and the matrix-core tiles), then row passes
    if (b->n_t1c > 0)
        hipLaunchKernelGGL(kb_t1_cols, dim3(b->n_t1c), dim3(256), 0, s, b->ddesc, b->own + b->off_t1c);
    NFFT2D_LAUNCH_CHECK();
    if (b->n_t1cm > 0)
        hipLaunchKernelGGL(kb_t1_cols_mm, dim3(b->n_t1cm), dim3(256), 0, s, b->ddesc, b->own + b->off_t1cm);
    NFFT2D_LAUNCH_CHECK();
    if (b->n_t1r > 0)
        hipLaunchKernelGGL(kb_t1_rows, dim3(b->n_t1r), dim3(128), 0, s, b->ddesc, b->own + b->off_t1r);
    NFFT2D_LAUNCH_CHECK();
    if (b->n_t1rm > 0)
        hipLaunchKernelGGL(kb_t1_rows_mm, dim3(b->n_t1rm), dim3(256), 0, s, b->ddesc, b->own + b->off_t1rm);
    NFFT2D_LAUNCH_CHECK();
    return NFFT2D_OK;
}

// setpts then type-1 (the spread needs the sorted points; one stream).
extern "C" int nfft2d_batch_setpts_type1(nfft2d_batch *b, void *stream)
{
    int rc = nfft2d_batch_setpts(b, stream);
    return rc ? rc : nfft2d_batch_type1(b, stream);
}

// setpts + type-2 of the batch with the bin sort on the batch's second stream,
// concurrent with the grid passes on `stream`; the gather waits for both.
extern "C" int nfft2d_batch_setpts_type2(nfft2d_batch *b, void *stream)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_setpts_type2: null batch");
    if (!b->have_points) return nfft2d_fail(NFFT2D_ERR_STATE, "batch_setpts_type2: nfft2d_batch_set_points was not called");
    if (!b->have_io) return nfft2d_fail(NFFT2D_ERR_STATE, "batch_setpts_type2: nfft2d_batch_set_io was not called");
    int rc = nfft2d_check_device(b->device);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = b->fk.init(b->device);
    if (rc || (rc = b->fk.fork(s))) return rc;
    rc = nfft2d_batch_setpts(b, b->fk.aux);
    const int rg = rc ? NFFT2D_OK : batch_grid(b, s);
    const int rj = b->fk.join(s);
    if (rc || (rc = rg) || (rc = rj)) return rc;
    return batch_gather(b, s);
}

// The batch's tuning, chosen by timing the whole batch (setpts + type-2),
// one parameter at a time (coordinate descent), the others held:
//   1. per size class (N1, N2): the subproblem size, 64, 128 or 256 points;
//   2. per size class: the direct-gather cut-off, 0, 4, 8 or 16 points
//      (subproblems up to it skip the LDS tile);
//   3. for the batch: the width of the persistent wave gather, 512, 1024 or
//      2048 blocks.
// None of them changes a result bit.  Synchronizes the stream; leaves the best
// values set.
extern "C" int nfft2d_batch_autotune(nfft2d_batch *b, void *stream, int reps)
{
    if (!b) return nfft2d_fail(NFFT2D_ERR_ARG, "batch_autotune: null batch");
    if (!b->have_points || !b->have_io)
        return nfft2d_fail(NFFT2D_ERR_STATE, "batch_autotune: set the points and the input/output arrays first");
    if (reps < 1) reps = 3;
    hipStream_t s = (hipStream_t)stream;
    static const int cand_mp[] = {64, 128, 256};
    static const int cand_dm[] = {0, 4, 8, 16};
    static const int cand_gw[] = {512, 1024, 2048};
    std::map<std::pair<int, int>, std::vector<int>> classes;
    for (int t = 0; t < b->count; ++t) classes[{b->hdesc[t].N1, b->hdesc[t].N2}].push_back(t);
    hipEvent_t e0, e1;
    NFFT2D_HIP_TRY(hipEventCreate(&e0));
    NFFT2D_HIP_TRY(hipEventCreate(&e1));
    int rc = NFFT2D_OK;
    // the faster of two timings of reps passes each, after a warm-up pass
    auto time_once = [&](float &ms) -> int {
        int r = nfft2d_batch_setpts_type2(b, s);
        if (r) return r;
        ms = 1e30f;
        for (int trial = 0; trial < 2; ++trial) {
            float t = 0.f;
            if (hipEventRecord(e0, s) != hipSuccess) return nfft2d_fail(NFFT2D_ERR_HIP, "batch_autotune: event record");
            for (int k = 0; k < reps && !r; ++k) r = nfft2d_batch_setpts_type2(b, s);
            if (r) return r;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                return nfft2d_fail(NFFT2D_ERR_HIP, "batch_autotune: event timing failed");
            ms = std::min(ms, t);
        }
        return NFFT2D_OK;
    };
    auto set_mp = [&](const std::vector<int> &ts, int mp) {
        for (int t : ts) { b->plans[t]->maxpts = mp; b->hdesc[t].maxpts = mp; }
        b->sorted = false;                        // the next setpts rebuilds the subproblem list
        return batch_upload(b);
    };
    auto set_dm = [&](const std::vector<int> &ts, int dm) {
        for (int t : ts) { b->dmax[t] = dm; b->hdesc[t].dmax = dm; }
        b->sorted = false;
        return batch_upload(b);
    };
    for (auto &kv : classes) {
        float best = 1e30f;
        int best_mp = b->hdesc[kv.second[0]].maxpts;
        for (int mp : cand_mp) {
            float ms = 0.f;
            if ((rc = set_mp(kv.second, mp)) || (rc = time_once(ms))) break;
            if (ms < best) { best = ms; best_mp = mp; }
        }
        if (rc || (rc = set_mp(kv.second, best_mp))) break;
    }
    for (auto &kv : classes) {
        if (rc) break;
        float best = 1e30f;
        int best_dm = b->dmax[kv.second[0]];
        for (int dm : cand_dm) {
            float ms = 0.f;
            if ((rc = set_dm(kv.second, dm)) || (rc = time_once(ms))) break;
            if (ms < best) { best = ms; best_dm = dm; }
        }
        if (rc || (rc = set_dm(kv.second, best_dm))) break;
    }
    if (!rc) {
        float best = 1e30f;
        int best_gw = b->gather_blocks;
        for (int gw : cand_gw) {
            float ms = 0.f;
            b->gather_blocks = gw;
            if ((rc = time_once(ms))) break;
            if (ms < best) { best = ms; best_gw = gw; }
        }
        b->gather_blocks = best_gw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
