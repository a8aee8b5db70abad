/*
 * nfft2d.h -- C ABI of the MI355X (gfx950) 2D NFFT library (libnfft2d.so).
 *
 * Drop-in replacement for the reference's 2D NFFT path
 * (OrangeOwlSolutions/Miscellaneous, paths relative to its root):
 *   NFFT/NFFT2_2D.cuh:4-11   NFFT2_2D_GPU (float2 / double2 overloads), the
 *                            global cuFFT plan and its create/destroy calls;
 *   NFFT/NDFT2_2D.cuh:4-6    NDFT2_2D_GPU;
 *   BP_FBP_FFBP/NFFT1_2D.m   the type-2 direction, which the reference has
 *                            only in MATLAB, exported here as NFFT1_2D_GPU.
 *
 * Conventions (the reference's, kept unchanged):
 *   - every pointer argument is DEVICE memory owned by the caller;
 *   - complex values are interleaved (re, im): nfft2d_c64 has the layout of
 *     float2 / cufftComplex, nfft2d_c128 that of double2 / double2_;
 *   - coordinates are in grid-index units: x has period N1, y has period N2
 *     (nominal range [-N/2, N/2)); any finite value is wrapped;
 *   - mode arrays are N1 x N2 row-major, mode (k1, k2), k in [-N/2, N/2), at
 *     index (k1 + N1/2) * N2 + (k2 + N2/2)  (NFFT/NFFT2_2D.cu:83);
 *   - transforms use e^{-2 pi i (k1 x / N1 + k2 y / N2)}, unnormalised;
 *   - c = 2 and K = 6 (NFFT/NFFT2_2D.cu:20-21);
 *   - result buffers are overwritten, never accumulated into.
 * Differences, all in the caller's favour:
 *   - every entry point returns a status (0 = NFFT2D_OK, < 0 = error; text in
 *     nfft2d_last_error()) instead of aborting the process;
 *   - no plan needs to exist before a drop-in call (plans are cached per
 *     size, precision and device), and the calls allocate nothing per call;
 *   - work is stream-ordered on the legacy default stream for the drop-ins,
 *     or on the caller's stream for the plan API; the host is not blocked.
 * Thread safety: the plan cache is internally locked.  Two calls that share
 * one plan on different streams race on the plan's scratch; calls on one
 * stream are ordered.
 * Devices: a plan (and a batch, an NDFT plan) belongs to the device current
 * at its creation; calls on it with another device current fail with
 * NFFT2D_ERR_ARG.  The drop-ins keep one cached plan per device.
 * Sizes: at most 2^30 points per transform (M), NFFT2D_ERR_ARG above.
 */
#ifndef NFFT2D_H
#define NFFT2D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y; } nfft2d_c64;
typedef struct { double x, y; } nfft2d_c128;
typedef struct nfft2d_plan_s nfft2d_plan;

enum {
    NFFT2D_OK = 0,
    NFFT2D_ERR_ARG = -1,          /* null pointer, negative count, bad flag   */
    NFFT2D_ERR_SIZE = -2,         /* N odd / non-positive / above the maximum */
    NFFT2D_ERR_HIP = -3,          /* a HIP runtime call failed                */
    NFFT2D_ERR_ALLOC = -4,        /* device allocation failed                 */
    NFFT2D_ERR_STATE = -5,        /* plan used before nfft2d_setpts           */
    NFFT2D_ERR_UNSUPPORTED = -6   /* K != 6 or c != 2                         */
};

enum { NFFT2D_C64 = 0, NFFT2D_C128 = 1 };

/* Last error of the calling thread, "" if none. */
const char *nfft2d_last_error(void);
/* Library version, major * 10000 + minor * 100 + patch. */
int nfft2d_version(void);

/* ------------------------------------------------------------------ drop-ins */

/* Type-1 (non-equispaced data -> equispaced result), the reference's
 * NFFT2_2D_GPU (NFFT/NFFT2_2D.cu:329 float, :413 double; decl.
 * NFFT/NFFT2_2D.cuh:8-11):
 *   result[k1][k2] = sum_{l<M} data[l] e^{-2 pi i (k1 x[l] / N1 + k2 y[l] / N2)}.
 * result: N1*N2, data: M, x, y: M.  N1, N2 even. */
int NFFT2_2D_GPU_c64(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y,
                     int N1, int N2, int M);
int NFFT2_2D_GPU_c128(nfft2d_c128 *result, const nfft2d_c128 *data, const double *x, const double *y,
                      int N1, int N2, int M);

/* Type-2 (equispaced data -> non-equispaced result), the reference's MATLAB
 * NFFT1_2D (BP_FBP_FFBP/NFFT1_2D.m:1, called as NFFT1_2D(data, x1, x2)):
 *   result[l] = sum_{k1,k2} data[k1][k2] e^{-2 pi i (k1 x[l] / N1 + k2 y[l] / N2)}.
 * result: M, data: N1*N2 row-major [k1][k2] (MATLAB's data(k1,k2)). */
int NFFT1_2D_GPU_c64(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y,
                     int N1, int N2, int M);
int NFFT1_2D_GPU_c128(nfft2d_c128 *result, const nfft2d_c128 *data, const double *x, const double *y,
                      int N1, int N2, int M);

/* NFFT2_2D_GPU_c64 that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as NFFT2_2D_GPU_c64 (the same kernels), plus the
 * stream the work is ordered on (NULL: the legacy default stream).  Every
 * buffer of the transform lives in one file-scope device array of 22 MiB that
 * each GPU the library runs on holds until the process ends; hence the limits
 * N1, N2 <= 256 (even) and M <= 262144, NFFT2D_ERR_SIZE beyond them.  All
 * calls on one GPU share that array: calls that overlap on one GPU (different
 * streams or host threads without ordering between them) race on it. */
int NFFT2_2D_GPU_c64_noalloc(nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, const float *y,
                             int N1, int N2, int M, void *stream);

/* Plan warm-up hooks, NFFT/NFFT2_2D.cu:49-50,55 (decl. NFFT/NFFT2_2D.cuh:5-7).
 * The reference builds one process-global cuFFT plan of size (2 N1) x (2 N2)
 * that the next NFFT2_2D_GPU call silently assumes; here they only pre-build
 * (or drop) the cached plans, and are optional. */
void Calculate_cuFFT_plan_C2C_NFFT2_2D(int N1, int N2);
void Calculate_cuFFT_plan_Z2Z_NFFT2_2D(int N1, int N2);
void Destroy_cuFFT_plan_NFFT2_2D(void);

/* Direct NED sum, NFFT/NDFT2_2D.cu:37 (decl. NFFT/NDFT2_2D.cuh:4-6):
 *   out[n] = sum_{m<M} in[m] e^{-2 pi i (u[n] X[m] / Nu + v[n] Y[m] / Nv)},  n < N,
 * in FP64.  `handle` took a cublasHandle_t; it is accepted and ignored.  The
 * reference materialises the M x N kernel matrix and writes it to a text file
 * (:44,56) before its gemv (:67); neither happens here.  Outputs on a meshgrid
 * (u, v arithmetic progressions, either flattening; detected on the device)
 * take the FP64 matrix-core contraction, any other (u, v) the direct sum.  Runs
 * on the legacy default stream.  A workspace of at most 64 N + 2 M complex
 * values is kept per device and grown when a call needs more (the growth
 * synchronises the device); concurrent calls on one device serialise on it. */
int NDFT2_2D_GPU(void *handle, const double *X, const double *Y, const double *u, const double *v,
                 const nfft2d_c128 *in, nfft2d_c128 *out, int Nu, int Nv, int M, int N);

/* No counterpart in the reference; for callers and tests that check which path
 * an NDFT2_2D_GPU call took.  Synchronises the current device and reports the
 * detection of its last NDFT2_2D_GPU call with N > 0: *is_grid = 1 when the
 * outputs were a grid of *nu x *nv values (*vmajor = 0: n = i nv + j, 1:
 * n = j nu + i) and took the matrix-core contraction, 0 when they took the
 * direct sum.  Before any such call: all zero. */
int nfft2d_ndft_grid_info(int *is_grid, int *nu, int *nv, int *vmajor);

/* ------------------------------------------------------------------ plan API */

/* A plan holds the twiddle and deconvolution tables, the oversampled grid and
 * the tile-sorted points.  dtype: NFFT2D_C64 or NFFT2D_C128.  K must be 6 and
 * c must be 2 (the reference's constants); flags must be 0.  N1, N2 even, at
 * most 4096 (complex64) or 2048 (complex128), any even value: powers of two
 * take pruned Stockham passes, other lengths from 128 on chirp-z (Bluestein)
 * passes, shorter ones direct O(N^2) passes. */
int nfft2d_plan_create(nfft2d_plan **plan, int N1, int N2, int dtype, int K, int c, unsigned flags);
int nfft2d_plan_destroy(nfft2d_plan *plan);

/* Tile-sort M points (x, y: float for C64 plans, double for C128 plans).  The
 * sorted copy lives in the plan; x and y may be freed afterwards. */
int nfft2d_setpts(nfft2d_plan *plan, const void *x, const void *y, int M, void *stream);

/* Type-2: result[l] = sum_k data[k] e^{-2 pi i k.x_l / N}  (result: M, data: N1*N2). */
int nfft2d_type2(nfft2d_plan *plan, void *result, const void *data, void *stream);
/* nfft2d_setpts followed by nfft2d_type2, with the tile sort on a second
 * stream owned by the plan, concurrent with the grid passes (pre-compensation,
 * zero padding, FFT) on `stream`; the gather waits for both.  Everything stays
 * ordered on `stream` for the caller and can be captured into a hipGraph (the
 * first call creates the second stream: make it outside a capture).  The
 * one-call form of the reference's type-2 (BP_FBP_FFBP/NFFT1_2D.m), which
 * NFFT1_2D_GPU_c64 / _c128 use. */
int nfft2d_setpts_type2(nfft2d_plan *plan, void *result, const void *data, const void *x, const void *y, int M,
                        void *stream);
/* Type-1: result[k] = sum_l data[l] e^{-2 pi i k.x_l / N}  (result: N1*N2, data: M). */
int nfft2d_type1(nfft2d_plan *plan, void *result, const void *data, void *stream);
/* Adjoint of type-2: result[k] = sum_l data[l] e^{+2 pi i k.x_l / N}. */
int nfft2d_adjoint(nfft2d_plan *plan, void *result, const void *data, void *stream);
/* Type-2 with the coordinate derivatives: y[l], dy1[l] = d y_l / d x_l,
 * dy2[l] = d y_l / d y_l-coordinate (all M complex). */
int nfft2d_type2_dx(nfft2d_plan *plan, void *y, void *dy1, void *dy2, const void *data, void *stream);
/* Gradient of a real loss with respect to the coordinates, given
 * g = dL/d(conj y) * 2 (PyTorch's convention for complex outputs):
 *   gx1[l] = Re(conj(g[l]) dy_l/dx_l),  gx2[l] likewise (M reals each).
 * Recomputes the oversampled grid from `data` (nothing is stored between the
 * forward and the backward). */
int nfft2d_type2_grad_x(nfft2d_plan *plan, void *gx1, void *gx2, const void *g, const void *data,
                        void *stream);
/* The same gradient on the oversampled grid the plan already holds: valid
 * right after nfft2d_type2 / nfft2d_type2_grid of the same data, before any
 * type-1 / adjoint / spread on the plan (those overwrite the grid).  Saves
 * the pre-compensation and FFT passes of the recomputation. */
int nfft2d_type2_grad_x_grid(nfft2d_plan *plan, void *gx1, void *gx2, const void *g, void *stream);
/* Least squares L = ||A(x) data - y_obs||^2 in one pass over the points:
 *   r[l] = 2 (y_l - y_obs[l])  (M complex; dL/d(conj y) times 2, the input of
 *          nfft2d_adjoint for dL/d data; r may be NULL),
 *   gx1[l], gx2[l] = Re(conj(r[l]) dy_l/dx_l), Re(conj(r[l]) dy_l/dy_l)  (M reals),
 * with the value and derivative sums sharing their tile reads.  Leaves the
 * oversampled grid of `data` in the plan (as nfft2d_type2 does), and r in the
 * plan, for nfft2d_lsq_adjoint.  The first call on a plan (and the first after
 * nfft2d_setpts with more points than before) allocates that residual buffer:
 * make it outside a graph capture. */
int nfft2d_type2_lsq(nfft2d_plan *plan, void *r, void *gx1, void *gx2, const void *y_obs, const void *data,
                     void *stream);
/* dL/d data of the same loss: the adjoint of the residual r that the last
 * nfft2d_type2_lsq left in the plan (result: N1*N2): nfft2d_adjoint(plan,
 * result, r), bit for bit when the oversampled grid is at least 44 x 44 (the
 * spread of smaller grids sums with LDS atomics, in an order that varies
 * between runs).  The spread reads the held residual in the plan's tile order
 * instead of gathering r by point index.
 * NFFT2D_ERR_STATE if no nfft2d_type2_lsq ran since the last nfft2d_setpts. */
int nfft2d_lsq_adjoint(nfft2d_plan *plan, void *result, void *stream);

/* The stages of the transforms, for callers that time them or reuse the
 * oversampled grid (which lives in the plan between the calls):
 *   type-2  = nfft2d_type2_grid + nfft2d_type2_interp
 *   type-1  = nfft2d_type1_spread + nfft2d_type1_finish(sign = -1)
 *   adjoint = nfft2d_type1_spread + nfft2d_type1_finish(sign = +1)
 * _grid: window pre-compensation, zero padding and the 2-D FFT
 *        (BP_FBP_FFBP/NFFT1_2D.m:48-55);
 * _interp: the (2K+1)^2 gather at the plan's points (NFFT1_2D.m:61-67);
 * _spread: zero the grid and spread the M samples onto it
 *          (NFFT/NFFT2_2D.cu:163-218);
 * _finish: 2-D FFT, crop and deconvolution (NFFT/NFFT2_2D.cu:380-402). */
int nfft2d_type2_grid(nfft2d_plan *plan, const void *data, void *stream);
int nfft2d_type2_interp(nfft2d_plan *plan, void *result, void *stream);
int nfft2d_type1_spread(nfft2d_plan *plan, const void *data, void *stream);
int nfft2d_type1_finish(nfft2d_plan *plan, void *result, int sign, void *stream);

/* Points per subproblem of the spread / gather: 32 to 256 (default 256). */
int nfft2d_plan_set_maxpts(nfft2d_plan *plan, int maxpts);
/* Number of points set on the plan; NFFT2D_ERR_STATE before the first
 * nfft2d_setpts (or after nfft2d_plan_set_maxpts), NFFT2D_ERR_ARG for a null
 * plan. */
int nfft2d_plan_num_points(const nfft2d_plan *plan);

/* ------------------------------------------------ NDFT as an MFMA contraction */

/* Type-2 direct sum over the equispaced modes, with no window and no grid:
 *   result[m] = sum_{k1,k2} data[k1][k2] e^{-2 pi i (k1 x[m] / N1 + k2 y[m] / N2)},
 * i.e. the sum of NFFT/NDFT2_2D.cu:15-32,67 and NFFT/NFFT2/Matlab/NDFT2_2D.m:9-14
 * in the modes -> points direction, evaluated as the separable contraction
 * y = rowsum(A o (B data^T)), A[m][k1] = e^{-2 pi i k1 x_m / N1},
 * B[m][k2] = e^{-2 pi i k2 y_m / N2}, on the matrix cores.  The product
 * B data^T runs in `precision`, always with FP32 accumulation:
 *   NFFT2D_NDFT_FP32    FP32 operands;
 *   NFFT2D_NDFT_BF16X3  each operand split x = hi + lo into two BF16 parts,
 *                       product hi*hi + hi*lo + lo*hi;
 *   NFFT2D_NDFT_FP16    FP16 operands, each k1 row of `data` scaled by a power
 *                       of two into the FP16 range (undone after the product).
 * The phasors A and B are evaluated in FP32 from an FP64 phase reduction.
 * Layouts and units are those of NFFT1_2D_GPU_c64.  The plan holds the mode
 * operand in the matrix cores' operand order (about 4 N1 N2 bytes for BF16X3);
 * nfft2d_ndft_set_modes fills it, nfft2d_ndft_eval contracts it with M
 * points, nfft2d_ndft_type2 does both.  N1, N2 even, at most 8192. */
typedef struct nfft2d_ndft_plan_s nfft2d_ndft_plan;
enum { NFFT2D_NDFT_FP32 = 0, NFFT2D_NDFT_BF16X3 = 1, NFFT2D_NDFT_FP16 = 2 };
int nfft2d_ndft_plan_create(nfft2d_ndft_plan **plan, int N1, int N2, int precision);
int nfft2d_ndft_plan_destroy(nfft2d_ndft_plan *plan);
int nfft2d_ndft_set_modes(nfft2d_ndft_plan *plan, const nfft2d_c64 *data, void *stream);
int nfft2d_ndft_eval(nfft2d_ndft_plan *plan, nfft2d_c64 *result, const float *x, const float *y, int M,
                     void *stream);
int nfft2d_ndft_type2(nfft2d_ndft_plan *plan, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x,
                      const float *y, int M, void *stream);

/* ------------------------------------------------------------------ 1-D lines */

/* Batched 1-D NFFTs, complex64, K = 6, c = 2: `batch` lines per call, each
 * line with N modes and M points, stored line after line (line b: modes at
 * b*N, points and samples at b*M).  N even, at most 2048, or a power of two up
 * to 4096 (type-1: at most 2048); other lengths from 128 on take chirp-z
 * (Bluestein) passes; coordinates in grid units with period N.
 *   nfft1d_type2  BP_FBP_FFBP/NFFT1_1D.m:1 (called once per pulse in
 *                 BP_FBP_FFBP/BP_NFFT.m:18):
 *                 result[b][l] = sum_k data[b][k + N/2] e^{-2 pi i k x[b][l] / N}
 *   nfft1d_type1  NFFT/NFFT2/1D/Matlab/NFFT2_1D.m:1 (NFFT2_1D(data, x, N, 2, 6),
 *                 Synthesis/Aper_Synth_1D/Aper_Synth_1D.m:74):
 *                 result[b][k + N/2] = sum_l data[b][l] e^{-2 pi i k x[b][l] / N}
 * The reference rounds c x half away from zero (MATLAB round); here half to
 * even, as the 2-D path (the two differ only at exact ties and are both valid
 * NFFTs).  The type-1 sums are exact in fixed point: its result does not
 * depend on the order of the additions.  The first call with more lines than
 * before allocates the plan's grid: make it outside a graph capture. */
typedef struct nfft1d_plan_s nfft1d_plan;
int nfft1d_plan_create(nfft1d_plan **plan, int N);
int nfft1d_plan_destroy(nfft1d_plan *plan);
int nfft1d_type2(nfft1d_plan *plan, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, int M, int batch,
                 void *stream);
int nfft1d_type1(nfft1d_plan *plan, nfft2d_c64 *result, const nfft2d_c64 *data, const float *x, int M, int batch,
                 void *stream);

/* ------------------------------------------------------------------ batches */

/* Many complex64 transforms of different sizes (K = 6, c = 2; N1, N2 even in
 * [2, 1024]: powers of two from 8 run Stockham passes, other sizes the
 * plan's chirp-z passes from 128 to 682 and direct ones otherwise, so that a
 * transform's result is bitwise its own plan's except for sizes from 684 on
 * that are not powers of two, where a plan runs chirp-z passes and the batch
 * direct ones and the two agree to rounding; any M), run as ONE kernel launch
 * per stage over all of them:
 * nfft2d_batch_setpts is 5 launches, nfft2d_batch_type2 and type1 4 each,
 * whatever the number of transforms.  Neither allocates nor records a
 * hipMemset, so both can be captured into one hipGraph and replayed.  There is
 * no reference counterpart: the reference runs one NFFT2_2D_GPU call per
 * transform (e.g. once per particle per generation in
 * Synthesis/PSO/PSO.cu:348).
 *   create       N1[t], N2[t], M[t] for t < count (host arrays); allocates
 *                every transform's plan and point buffers;
 *   set_points   x[t], y[t]: device coordinate arrays (host array of device
 *                pointers), kept by reference;
 *   set_io       modes[t] (N1 x N2, layout of NFFT1_2D_GPU_c64) and result[t]
 *                (M[t]): device arrays, kept by reference;
 *   setpts       tile-sort the points of every transform;
 *   type2        result[t] = type-2 of modes[t] at the sorted points;
 *   setpts_type2 setpts + type2, the tile sort on a second stream owned by the
 *                batch, concurrent with the grid passes (as
 *                nfft2d_setpts_type2; the first call creates the stream);
 *   autotune     choose, by timing reps setpts_type2 passes of the whole
 *                batch per candidate, one parameter at a time: per size class
 *                (N1, N2) the subproblem size (64, 128 or 256 points), then per
 *                size class the direct-gather cut-off (0, 4, 8 or 16 points),
 *                then the width of the persistent gather (512, 1024 or 2048
 *                blocks); no choice changes a result bit; synchronizes the
 *                stream;
 *   set_maxpts / get_maxpts  the subproblem size of transform t ([64, 256]);
 *   set_direct / get_direct  transform t's direct-gather cut-off ([0, 16]):
 *                its subproblems of at most that many points are gathered
 *                point by point from the grid instead of through an LDS tile
 *                (default 8); like set_maxpts, it takes effect at the next
 *                setpts;
 *   set_gather_blocks / get_gather_blocks  blocks of the persistent gather
 *                (default 1024). */
typedef struct nfft2d_batch_s nfft2d_batch;
int nfft2d_batch_create(nfft2d_batch **batch, int count, const int *N1, const int *N2, const int *M, unsigned flags);
int nfft2d_batch_destroy(nfft2d_batch *batch);
int nfft2d_batch_count(const nfft2d_batch *batch);
int nfft2d_batch_set_points(nfft2d_batch *batch, const float *const *x, const float *const *y);
int nfft2d_batch_set_io(nfft2d_batch *batch, const nfft2d_c64 *const *modes, nfft2d_c64 *const *result);
int nfft2d_batch_setpts(nfft2d_batch *batch, void *stream);
int nfft2d_batch_type2(nfft2d_batch *batch, void *stream);
int nfft2d_batch_setpts_type2(nfft2d_batch *batch, void *stream);
int nfft2d_batch_autotune(nfft2d_batch *batch, void *stream, int reps);
int nfft2d_batch_set_maxpts(nfft2d_batch *batch, int t, int maxpts);
int nfft2d_batch_get_maxpts(const nfft2d_batch *batch, int t);
int nfft2d_batch_set_direct(nfft2d_batch *batch, int t, int npts);
int nfft2d_batch_get_direct(const nfft2d_batch *batch, int t);
int nfft2d_batch_set_gather_blocks(nfft2d_batch *batch, int blocks);
int nfft2d_batch_get_gather_blocks(const nfft2d_batch *batch);
/* Type-1 of every transform (the reference's NFFT2_2D_GPU, NFFT/NFFT2_2D.cu:329,
 * once per transform; e.g. once per particle of a PSO generation,
 * Synthesis/PSO/PSO.cu:348 -> Synthesis.cu:395): result[t][k1][k2] =
 * sum_l data[t][l] e^{-2 pi i (k1 x[t][l] / N1 + k2 y[t][l] / N2)}.
 *   set_io_type1   data[t] (M[t] samples) and result[t] (N1 x N2): device
 *                  arrays, kept by reference; allocates each transform's
 *                  spread tiles; N1, N2 >= 22 (NFFT2D_ERR_SIZE otherwise);
 *   type1          4 launches (spread, tile merge, the two FFT passes with
 *                  crop and deconvolution), needs setpts;
 *   setpts_type1   setpts then type1 on `stream` (9 launches). */
int nfft2d_batch_set_io_type1(nfft2d_batch *batch, const nfft2d_c64 *const *data, nfft2d_c64 *const *result);
int nfft2d_batch_type1(nfft2d_batch *batch, void *stream);
int nfft2d_batch_setpts_type1(nfft2d_batch *batch, void *stream);

/* ------------------------------------------------------------------------
 * Type-3 2-D NUFFT (non-uniform to non-uniform), FP64:
 *   F[j] = sum_k f[k] e^{-i (s[j] x[k] + t[j] y[k])},  j < len_out, k < len_in,
 * to the accuracy eps, by Gaussian gridding with the parameters of
 * NUFFT_2D_CPU (NFFT/NFFT3/C/NFFT3_2D_C.cpp:199, declared at
 * NFFT/NFFT3/C/NFFT3_2D_C.h:6-15): R = 2.2, 2 msp + 1 taps per dimension from
 * eps, an Mry x Mrx grid from the extents of x, y, s, t.  Same argument order
 * and meaning; f, F as fftw_complex (interleaved re, im).
 *   NUFFT_2D_GPU         device arrays; runs on the null stream, synchronises
 *                        once (the extents go to the host to size the grid);
 *   NUFFT_2D_GPU_host    host arrays, the contract of NUFFT_2D_CPU itself
 *                        (copies in, runs, copies out; returns when F is set);
 *   NUFFT_2D_GPU_params  the 14 derived values, in the order
 *                        xb, yb, sb, tb, kk_x, kk_y, Dx, Dy, b, t1, fact^2, msp,
 *                        Mrx, Mry (:214-267), computed from device arrays.
 * Errors (NFFT2D_ERR_ARG): eps outside (0, 1), a non-finite coordinate, all x
 * or all y equal (the reference divides by the half-extent there).
 * NFFT2D_ERR_SIZE: msp > 42 (eps below ~1e-14) or Mrx, Mry above 4096
 * (8192 for powers of two).  len_in = 0 gives F = 0; len_out = 0 does
 * nothing.  Not bitwise reproducible across runs (inputs are binned in
 * arrival order); calls on one device serialise on a shared workspace. */
int NUFFT_2D_GPU(const double *x, const double *y, const double *s, const double *t, const nfft2d_c128 *f,
                 nfft2d_c128 *F, double eps, int len_in, int len_out);
int NUFFT_2D_GPU_host(const double *x, const double *y, const double *s, const double *t, const nfft2d_c128 *f,
                      nfft2d_c128 *F, double eps, int len_in, int len_out);
int NUFFT_2D_GPU_params(const double *x, const double *y, const double *s, const double *t, double eps, int len_in,
                        int len_out, double *params);

/* ------------------------------------------------------------------------
 * Band-limited polar interpolation, the steps around its type-2 NFFT
 * (BP_FBP_FFBP/knabNFFT2D.m:1-79 and rectNFFT2D.m:1-77, MATLAB functions with
 * no C interface in the reference; the Python knabNFFT2D / rectNFFT2D call
 * these with the 2-D FFT of the samples and a type-2 plan in between).
 * Device arrays, FP64 coordinates, launched on `stream`.  rho_out and cos_out
 * are R x C; out_strides (host, 4 entries, in elements: rho_out's row and
 * column strides, then cos_out's) lets meshgrid views in without a copy, null
 * means both contiguous.
 *   nfft2d_knab_stats     out[0..9] = mean step of the rho axis and of the
 *                         cos alpha axis (mean of the differences), their first
 *                         values, two zeros, then min, max of the rho axis and
 *                         of the cos alpha axis (knabNFFT2D.m:28-40, :76-77);
 *                         out[NFFT2D_KNAB_STATS + 4 b ...] = min, max of rho_out
 *                         and min, max of cos_out over block b's share, for b <
 *                         NFFT2D_KNAB_STAT_BLOCKS (the caller folds them into
 *                         the extents).  The axes are read with element
 *                         strides.  NaN propagates.
 *   nfft2d_knab_spectrum  F (PN x QM, dtype) from the caller's 2-D spectrum
 *                         S (M x N, cos alpha x rho): the band of 2 kgn + 1 by
 *                         2 kgm + 1 frequencies centred at (PN/2, QM/2) holds
 *                         conj(S[(v - kgm) mod M][(u - kgn) mod N] w), w =
 *                         Gn[u] Gm[v] scale (FP64, rounded once), or w = scale
 *                         when Gn and Gm are null; zero elsewhere
 *                         (knabNFFT2D.m:58-70, the window of KnabWindow.m).
 *   nfft2d_knab_points    x1, x2 (R2 x C2, float for NFFT2D_C64, double for
 *                         NFFT2D_C128) = wrap((rho_out - r0) / h1, P1) and
 *                         wrap((cos_out - c0) / h2, P2), wrap(x, P) = x -
 *                         P rint(x / P) in FP64; entries beyond the caller's
 *                         R x C take 0 (the padded branch, knabNFFT2D.m:45-56).
 *   nfft2d_knab_finish    result (R x C) = conj(y[r C2 + c]) / scale, or
 *                         extrap where rho_out[0][c] is outside [rho_min,
 *                         rho_max] or cos_out[r][0] outside [cos_min, cos_max]
 *                         (knabNFFT2D.m:74-79). */
enum { NFFT2D_KNAB_STATS = 10, NFFT2D_KNAB_STAT_BLOCKS = 256 };
int nfft2d_knab_stats(const double *rho_axis, int n_rho, long long rho_stride, const double *cos_axis, int n_cos,
                      long long cos_stride, const double *rho_out, const double *cos_out, int R, int C,
                      const long long *out_strides, double *out, void *stream);
int nfft2d_knab_spectrum(int dtype, const void *S, int M, int N, int kgn, int kgm, const double *Gn, const double *Gm,
                         double scale, void *F, int PN, int QM, void *stream);
int nfft2d_knab_points(int dtype, const double *rho_out, const double *cos_out, int R, int C,
                       const long long *out_strides, int R2, int C2, double r0, double c0, double h1, double h2,
                       double P1, double P2, void *x1, void *x2, void *stream);
int nfft2d_knab_finish(int dtype, const void *y, int R, int C, int C2, double scale, const double *rho_out,
                       const double *cos_out, const long long *out_strides, double rho_min, double rho_max,
                       double cos_min, double cos_max, double extrap_re, double extrap_im, void *result, void *stream);

/* ------------------------------------------------------------------------
 * Testing hooks (no counterpart in the reference).  Every device allocation of
 * the library is numbered in order; nfft2d_debug_fail_alloc(n) makes the n-th
 * allocation from now (0 = the next) fail as if the device were out of
 * memory, once; n < 0 disables it.  nfft2d_debug_alloc_count() returns how
 * many allocations were attempted so far.  For tests of the error paths. */
void nfft2d_debug_fail_alloc(long long n);
long long nfft2d_debug_alloc_count(void);

#ifdef __cplusplus
}
#endif

#endif /* NFFT2D_H */
