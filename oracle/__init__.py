"""CPU oracle for the 2D NFFT path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``miscellaneous_amd``) never imports it.

It wraps ``liboracle_nfft2d.so`` (built from ``nfft2d_oracle.c`` by
``oracle/Makefile``), a plain-C step-by-step restatement of the reference
algorithm:

* ``nfft1_2d``  -- type-2 (uniform -> non-uniform), ``BP_FBP_FFBP/NFFT1_2D.m``.
* ``nfft2_2d``  -- type-1 (non-uniform -> uniform), ``NFFT/NFFT2_2D.cu:329-495``
  and its MATLAB twin ``NFFT/NFFT2/Matlab/NFFT2_2D.m``.
* ``ndft2_2d``  -- the direct NED sum, ``NFFT/NDFT2_2D.cu:28-29,67``.
* ``nfft1_1d`` / ``nfft2_1d`` -- the 1-D type-2 (``BP_FBP_FFBP/NFFT1_1D.m``) and type-1
  (``NFFT/NFFT2/1D/Matlab/NFFT2_1D.m``), with their direct sums ``ndft1d_type2`` /
  ``ndft1d_type1`` (the closed forms of ``NFFT1_1D_BLAS.m`` / ``NFFT2_1D_BLAS.m``).
* ``ndft_type2`` / ``ndft_type1`` -- direct FP64 sums of both directions (the
  closed forms every fast result is checked against).
* ``window``, ``window_deriv``, ``bessel_i0``, ``spectral_window``, ``fft2`` --
  the pieces, exposed so that each can be pinned on its own.
* ``nufft3_2d`` -- the type-3 (non-uniform -> non-uniform) Gaussian-gridding
  NUFFT, ``NFFT/NFFT3/C/NFFT3_2D_C.cpp:199-312`` (``NUFFT_2D_CPU``), with
  ``nufft3_params`` (its parameters, :214-267) and the direct FP64 sum
  ``ndft3_2d``.
* ``oracle.knab`` (its own module) -- the band-limited polar interpolation
  ``BP_FBP_FFBP/knabNFFT2D.m`` and its helpers, on ``nfft1_2d``.

Pins (what checks each function, in ``tests/test_oracle.py``):

* ``bessel_i0``: ``scipy.special.i0`` and 50-digit mpmath values.
* ``window``: 50-digit mpmath values of the closed form, and the Fourier pair
  phi_hat(d) = (1/2pi) int_{-alpha}^{alpha} I0(K sqrt(alpha^2-xi^2)) e^{i xi d} dxi
  by quadrature.
* ``window_deriv``: central differences of the mpmath closed form.
* ``fft2``: ``numpy.fft.fft2`` and impulse / single-exponential cases.
* ``nfft1_2d`` / ``nfft2_2d``: committed golden vectors from FP64 direct sums
  (numpy, ``tests/golden/make_golden.py``), the adjoint identity, the shift
  theorem, periodicity, single-point and single-mode closed forms, and the
  reference's own relation NDFT2_2D(data,x,y,N,M) ~ NFFT2_2D(data,y,x,M,N)
  (``NFFT/NFFT2/Matlab/Test_NFFT2_2D.m:24-29``).
* ``ndft2_2d`` / ``ndft_type2`` / ``ndft_type1``: brute-force numpy sums
  written independently in the golden script, and each other.
* ``ndft_type1_modes``: the golden type-1 sums at seeded modes (square and
  non-square, where swapped indices or periods fail).
* ``ndft1d_type2`` / ``ndft1d_type1``: numpy brute-force sums of the BLAS
  forms; ``nfft1_1d`` / ``nfft2_1d``: those direct sums (the algorithm's
  error at c = 2, K = 6 is ~1e-12), the adjoint identity between them, and
  the 2-D restatement on one mode row.
* ``ndft3_2d``: a numpy brute-force sum.  ``nufft3_2d``: that direct sum at
  eps = 1e-6, 1e-9, 1e-12 (the Gaussian gridding's error is ~eps; each of three
  planted mistakes -- no checkerboard sign, no post-compensation, flipped
  pre-modulation -- fails it), anisotropic and off-centre extents, and
  single-input closed forms.  ``nufft3_params``: the grid is even and at least
  the bound of :237-238, and the scaled extents are +-pi/2.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_nfft2d.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_alpha.restype = ctypes.c_double
    lib.orc_alpha.argtypes = [ctypes.c_int]
    lib.orc_bessel_i0.restype = ctypes.c_double
    lib.orc_bessel_i0.argtypes = [ctypes.c_double]
    for name in ("orc_window", "orc_window_series", "orc_window_deriv"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_double
        fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.orc_spectral_window.restype = None
    lib.orc_spectral_window.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for sfx, p in (("_d", _dp), ("_f", _fp)):
        fn = getattr(lib, "orc_fft2" + sfx)
        fn.restype = None
        fn.argtypes = [p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("orc_nfft1_2d", "orc_nfft2_2d"):
            fn = getattr(lib, name + sfx)
            fn.restype = ctypes.c_int
            fn.argtypes = [p, p, p, p] + [ctypes.c_int] * 7
    lib.orc_ndft2_2d.restype = None
    lib.orc_ndft2_2d.argtypes = [_dp] * 6 + [ctypes.c_int] * 4
    lib.orc_ndft_type2.restype = None
    lib.orc_ndft_type2.argtypes = [_dp] * 6 + [ctypes.c_int] * 3
    lib.orc_ndft_type1.restype = None
    lib.orc_ndft_type1.argtypes = [_dp] * 4 + [ctypes.c_int] * 3
    lib.orc_ndft_type1_modes.restype = None
    lib.orc_ndft_type1_modes.argtypes = [_dp] * 4 + [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] * 2 + [ctypes.c_int]
    for name in ("orc_nfft1_1d", "orc_nfft2_1d"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [_dp] * 3 + [ctypes.c_int] * 5
    for name in ("orc_ndft1d_type2", "orc_ndft1d_type1"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [_dp] * 3 + [ctypes.c_int] * 2
    lib.orc_nufft3_2d.restype = None
    lib.orc_nufft3_2d.argtypes = [_dp] * 6 + [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                              ctypes.c_int]
    lib.orc_nufft3_params.restype = None
    lib.orc_nufft3_params.argtypes = [_dp] * 4 + [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_ndft3_2d.restype = None
    lib.orc_ndft3_2d.argtypes = [_dp] * 6 + [ctypes.c_int, ctypes.c_int]
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_set_num_threads.argtypes = [ctypes.c_int]
    _lib = lib
    return lib


def _ptr(a: np.ndarray, real):
    return a.ctypes.data_as(ctypes.POINTER(real))


def _cplx_in(a, dtype):
    """Complex array -> contiguous interleaved real view of the given precision."""
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.view(np.float64 if dtype == np.complex128 else np.float32)


# ----------------------------------------------------------------------------- scalars
def alpha(c: int = 2) -> float:
    return _load().orc_alpha(c)


def bessel_i0(x: float) -> float:
    return _load().orc_bessel_i0(float(x))


def window(d, K: int = 6, c: int = 2, series: bool = False):
    lib = _load()
    fn = lib.orc_window_series if series else lib.orc_window
    a = alpha(c)
    return np.array([fn(float(v), K, a) for v in np.atleast_1d(d)])


def window_deriv(d, K: int = 6, c: int = 2):
    lib = _load()
    a = alpha(c)
    return np.array([lib.orc_window_deriv(float(v), K, a) for v in np.atleast_1d(d)])


def spectral_window(N: int, K: int = 6, c: int = 2) -> np.ndarray:
    out = np.zeros(N)
    _load().orc_spectral_window(_ptr(out, ctypes.c_double), N, K, c)
    return out


def num_threads() -> int:
    return _load().orc_num_threads()


def set_num_threads(n: int) -> None:
    _load().orc_set_num_threads(int(n))


# ----------------------------------------------------------------------------- transforms
def fft2(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised 2-D DFT (sign -1 = forward e^{-i})."""
    dt = np.complex128 if a.dtype != np.complex64 else np.complex64
    a, v = _cplx_in(a.copy(), dt)
    real = ctypes.c_double if dt == np.complex128 else ctypes.c_float
    fn = _load().orc_fft2_d if dt == np.complex128 else _load().orc_fft2_f
    fn(_ptr(v, real), a.shape[0], a.shape[1], sign)
    return a


def nfft1_2d(data: np.ndarray, x1: np.ndarray, x2: np.ndarray, K: int = 6, c: int = 2,
             round_mode: int = 0, mutation: int = 0, precision: str = "double") -> np.ndarray:
    """Type-2 NFFT (BP_FBP_FFBP/NFFT1_2D.m). data: (N1, N2) centred; x1, x2: grid units."""
    dt = np.complex128 if precision == "double" else np.complex64
    rt = np.float64 if precision == "double" else np.float32
    real = ctypes.c_double if precision == "double" else ctypes.c_float
    N1, N2 = data.shape
    M = len(x1)
    d, dv = _cplx_in(data, dt)
    x1 = np.ascontiguousarray(x1, dtype=rt)
    x2 = np.ascontiguousarray(x2, dtype=rt)
    out = np.zeros(M, dtype=dt)
    ov = out.view(rt)
    fn = _load().orc_nfft1_2d_d if precision == "double" else _load().orc_nfft1_2d_f
    rc = fn(_ptr(ov, real), _ptr(dv, real), _ptr(x1, real), _ptr(x2, real), N1, N2, M, K, c,
            round_mode, mutation)
    if rc != 0:
        raise ValueError("oracle nfft1_2d: unsupported arguments")
    return out


def nfft2_2d(data: np.ndarray, x: np.ndarray, y: np.ndarray, N1: int, N2: int, K: int = 6,
             c: int = 2, round_mode: int = 0, mutation: int = 0,
             precision: str = "double") -> np.ndarray:
    """Type-1 NFFT (NFFT/NFFT2_2D.cu NFFT2_2D_GPU).  Returns (N1, N2) centred."""
    dt = np.complex128 if precision == "double" else np.complex64
    rt = np.float64 if precision == "double" else np.float32
    real = ctypes.c_double if precision == "double" else ctypes.c_float
    M = len(x)
    d, dv = _cplx_in(data, dt)
    x = np.ascontiguousarray(x, dtype=rt)
    y = np.ascontiguousarray(y, dtype=rt)
    out = np.zeros((N1, N2), dtype=dt)
    ov = out.view(rt)
    fn = _load().orc_nfft2_2d_d if precision == "double" else _load().orc_nfft2_2d_f
    rc = fn(_ptr(ov, real), _ptr(dv, real), _ptr(x, real), _ptr(y, real), N1, N2, M, K, c,
            round_mode, mutation)
    if rc != 0:
        raise ValueError("oracle nfft2_2d: unsupported arguments")
    return out


def ndft2_2d(X, Y, u, v, data, Nu: int, Nv: int) -> np.ndarray:
    """Direct NED sum of NFFT/NDFT2_2D.cu: out[n] = sum_m in[m] e^{-2pi i(u_n X_m/Nu + v_n Y_m/Nv)}."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    d, dv = _cplx_in(data, np.complex128)
    N = len(u)
    out = np.zeros(N, dtype=np.complex128)
    _load().orc_ndft2_2d(_ptr(X, ctypes.c_double), _ptr(Y, ctypes.c_double),
                         _ptr(u, ctypes.c_double), _ptr(v, ctypes.c_double),
                         _ptr(dv, ctypes.c_double), _ptr(out.view(np.float64), ctypes.c_double),
                         Nu, Nv, len(X), N)
    return out


def ndft_type2(f: np.ndarray, x1, x2, with_dx: bool = False):
    """Direct FP64 type-2 sum (and optionally dy/dx1, dy/dx2)."""
    N1, N2 = f.shape
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    M = len(x1)
    fc, fv = _cplx_in(f, np.complex128)
    y = np.zeros(M, dtype=np.complex128)
    dx1 = np.zeros(M, dtype=np.complex128) if with_dx else None
    dx2 = np.zeros(M, dtype=np.complex128) if with_dx else None
    null = ctypes.POINTER(ctypes.c_double)()
    _load().orc_ndft_type2(_ptr(y.view(np.float64), ctypes.c_double),
                           _ptr(dx1.view(np.float64), ctypes.c_double) if with_dx else null,
                           _ptr(dx2.view(np.float64), ctypes.c_double) if with_dx else null,
                           _ptr(fv, ctypes.c_double), _ptr(x1, ctypes.c_double),
                           _ptr(x2, ctypes.c_double), N1, N2, M)
    return (y, dx1, dx2) if with_dx else y


def ndft_type1(d: np.ndarray, x1, x2, N1: int, N2: int) -> np.ndarray:
    """Direct FP64 type-1 sum, (N1, N2) centred output."""
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    dc, dv = _cplx_in(d, np.complex128)
    F = np.zeros((N1, N2), dtype=np.complex128)
    _load().orc_ndft_type1(_ptr(F.view(np.float64), ctypes.c_double), _ptr(dv, ctypes.c_double),
                           _ptr(x1, ctypes.c_double), _ptr(x2, ctypes.c_double), N1, N2, len(x1))
    return F


def ndft_type1_modes(d: np.ndarray, x1, x2, N1: int, N2: int, i1, i2) -> np.ndarray:
    """Direct FP64 type-1 sum at the centred mode indices (i1[q], i2[q])."""
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    dc, dv = _cplx_in(d, np.complex128)
    i1 = np.ascontiguousarray(i1, dtype=np.int32)
    i2 = np.ascontiguousarray(i2, dtype=np.int32)
    F = np.zeros(len(i1), dtype=np.complex128)
    _load().orc_ndft_type1_modes(_ptr(F.view(np.float64), ctypes.c_double), _ptr(dv, ctypes.c_double),
                                 _ptr(x1, ctypes.c_double), _ptr(x2, ctypes.c_double), N1, N2, len(x1),
                                 _ptr(i1, ctypes.c_int), _ptr(i2, ctypes.c_int), len(i1))
    return F


def _line(a, dtype=np.complex128):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.view(np.float64)


def nfft1_1d(data: np.ndarray, x, K: int = 6, c: int = 2, round_mode: int = 1) -> np.ndarray:
    """Type-2 1-D NFFT, BP_FBP_FFBP/NFFT1_1D.m: y_l ~ sum_k data_k e^{-2 pi i k x_l / N}
    (data: N modes, k at index k + N/2; x in grid units).  round_mode 1 = MATLAB round."""
    d, dv = _line(data)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.complex128)
    _load().orc_nfft1_1d(_ptr(out.view(np.float64), ctypes.c_double), _ptr(dv, ctypes.c_double),
                         _ptr(x, ctypes.c_double), len(d), len(x), K, c, round_mode)
    return out


def nfft2_1d(data: np.ndarray, x, N: int, K: int = 6, c: int = 2, round_mode: int = 1) -> np.ndarray:
    """Type-1 1-D NFFT, NFFT/NFFT2/1D/Matlab/NFFT2_1D.m (called there as NFFT2_1D(data, x, N, c, K)):
    F_k ~ sum_l data_l e^{-2 pi i k x_l / N}, k in [-N/2, N/2) at index k + N/2."""
    d, dv = _line(data)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(N, dtype=np.complex128)
    _load().orc_nfft2_1d(_ptr(out.view(np.float64), ctypes.c_double), _ptr(dv, ctypes.c_double),
                         _ptr(x, ctypes.c_double), N, len(x), K, c, round_mode)
    return out


def ndft1d_type2(f: np.ndarray, x) -> np.ndarray:
    """Direct FP64 1-D type-2 sum (NFFT/NFFT1/1D/Matlab/NFFT1_1D_BLAS.m in grid units)."""
    fd, fv = _line(f)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.complex128)
    _load().orc_ndft1d_type2(_ptr(out.view(np.float64), ctypes.c_double), _ptr(fv, ctypes.c_double),
                             _ptr(x, ctypes.c_double), len(fd), len(x))
    return out


def ndft1d_type1(d: np.ndarray, x, N: int) -> np.ndarray:
    """Direct FP64 1-D type-1 sum (NFFT/NFFT2/1D/Matlab/NFFT2_1D_BLAS.m in grid units)."""
    dd, dv = _line(d)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(N, dtype=np.complex128)
    _load().orc_ndft1d_type1(_ptr(out.view(np.float64), ctypes.c_double), _ptr(dv, ctypes.c_double),
                             _ptr(x, ctypes.c_double), N, len(x))
    return out


def _reals(*arrs):
    return [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]


def nufft3_2d(x, y, s, t, f, eps: float, mutation: int = 0):
    """Type-3 2-D NUFFT, NFFT/NFFT3/C/NFFT3_2D_C.cpp:199-312 (NUFFT_2D_CPU) restated step by step:
    F_j ~ sum_k f_k e^{-i (s_j x_k + t_j y_k)}.  Returns (F, (Mrx, Mry)).  mutation != 0 plants a
    known mistake (tests only): 1 no checkerboard sign, 2 no post-compensation, 3 flipped
    pre-modulation phase."""
    x, y, s, t = _reals(x, y, s, t)
    fd, fv = _line(f)
    F = np.zeros(len(s), dtype=np.complex128)
    g = (ctypes.c_int * 2)()
    _load().orc_nufft3_2d(_ptr(F.view(np.float64), ctypes.c_double), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
                          _ptr(s, ctypes.c_double), _ptr(t, ctypes.c_double), _ptr(fv, ctypes.c_double), float(eps),
                          len(x), len(s), g, int(mutation))
    return F, (g[0], g[1])


def nufft3_params(x, y, s, t, eps: float) -> dict:
    """The parameters NUFFT_2D_CPU derives from the data extents (NFFT3_2D_C.cpp:214-267)."""
    x, y, s, t = _reals(x, y, s, t)
    out = np.zeros(14)
    _load().orc_nufft3_params(_ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), _ptr(s, ctypes.c_double),
                              _ptr(t, ctypes.c_double), float(eps), len(x), len(s), _ptr(out, ctypes.c_double))
    keys = ["xb", "yb", "sb", "tb", "kkx", "kky", "Dx", "Dy", "b", "t1", "fact", "msp", "Mrx", "Mry"]
    return {k: (int(v) if k in ("msp", "Mrx", "Mry") else float(v)) for k, v in zip(keys, out)}


def ndft3_2d(x, y, s, t, f) -> np.ndarray:
    """Direct FP64 type-3 sum F_j = sum_k f_k e^{-i (s_j x_k + t_j y_k)}."""
    x, y, s, t = _reals(x, y, s, t)
    fd, fv = _line(f)
    F = np.zeros(len(s), dtype=np.complex128)
    _load().orc_ndft3_2d(_ptr(F.view(np.float64), ctypes.c_double), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
                         _ptr(s, ctypes.c_double), _ptr(t, ctypes.c_double), _ptr(fv, ctypes.c_double), len(x), len(s))
    return F


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
