"""Band-limited 2-D interpolation of polar-grid samples by the type-2 NFFT.

The reference's ``knabNFFT2D`` (``BP_FBP_FFBP/knabNFFT2D.m:1-79``), the step
above ``NFFT1_2D`` in its filtered and fast-factorised backprojection
(``BP_FBP_FFBP/FBP.m:95``, ``FFBP_POLAR_NFFT.m:89,141``): samples on a regular
(rho, cos alpha) grid are interpolated at arbitrary output points with
J. Selva's convolution-based trigonometric interpolation
(``NFFT_Based_Selvas_Approach_for_Signal_Interp_2D.m:4-28``):

    sF  = fft2(samples), repeated periodically to 2 kg + 1 frequencies per
          dimension and multiplied by the window G = Gn (x) Gm
          (``KnabWindow.m``: a pass band of 1/N, tapered by the convolution of
          a rectangle with an approximate prolate window, ``ConvWindow.m``);
    pad sF with zeros to a N x b M (oversampling a, b detected from the
          output sampling density);
    out = conj(type2(conj(sF), rho / (T1 / a), cos / (T2 / b))) / (a b N M),

then output points outside the input grid get ``ExtrapVal``.  ``rectNFFT2D``
(``BP_FBP_FFBP/rectNFFT2D.m:1-77``) is the same with a rectangular window.

Here the type-2 runs on the library (``Plan.setpts_type2``: tile sort, grid
passes and gather; complex64, or complex128 for complex128 samples), and the
steps around it are the library's ``nfft2d_knab_*`` kernels: one reduction for
the grid steps, origins and extents (one transfer to the host, which sizes the
transform), one pass that windows, pads and conjugates the spectrum, one that
scales and wraps the output points, one that conjugates, scales, crops and
marks points outside the input grid.  The input spectrum is a plain 2-D FFT of
the caller's samples (``torch.fft.fft2``, as the reference calls MATLAB's
``fft2``), and the window tables are computed on the host in FP64 once per
(T, B, P, N): the reference integrates ``ConvWindow`` with MATLAB's adaptive
``quad``; here the substitution f = B/2 sin(theta) removes the square-root
edge of the integrand and a 64-point Gauss-Legendre rule integrates the
remaining smooth function to FP64 rounding.

Sizes: a N and b M must be even (the library's even-size plans; the reference
also has an odd branch), and at most 4096 (complex64) or 2048 (complex128).

Repeated calls on one geometry (the same grid tensor objects, not modified in
place since, on the same stream: the version counters tell) reuse the grid
statistics, the output points and the plan's tile sort: such a call does not
synchronise with the host, and costs the spectrum, its windowing, the type-2's
grid passes and gather, and the final pass.
"""
from __future__ import annotations

import ctypes
import functools
import math

import numpy as np
import torch

from . import _lib
from ._lib import NFFT2D_C64, NFFT2D_C128, NFFT2D_KNAB_STAT_BLOCKS, NFFT2D_KNAB_STATS, check
from .nfft import Plan, _ptr, _stream

__all__ = ["knabNFFT2D", "rectNFFT2D", "knab_window"]

_GL_X, _GL_W = np.polynomial.legendre.leggauss(64)


def _conv_window_integral(lo: float, hi: float, B: float, T: float) -> float:
    """int_lo^hi W(f) df for W(f) = I0(pi B T/2 sqrt(1 - (2f/B)^2)) / (B sinh(pi B T/2)/(pi B T/2)),
    [lo, hi] inside [-B/2, B/2]; with f = B/2 sin(theta): W df = I0(a cos theta) cos theta B/2 dtheta."""
    a = math.pi * B * T / 2
    t0 = math.asin(max(-1.0, min(1.0, 2 * lo / B)))
    t1 = math.asin(max(-1.0, min(1.0, 2 * hi / B)))
    th = 0.5 * (t1 - t0) * _GL_X + 0.5 * (t1 + t0)
    vals = np.i0(a * np.cos(th)) * np.cos(th)
    return float(0.5 * (t1 - t0) * np.dot(_GL_W, vals) * (B / 2) / (B * math.sinh(a) / a))


@functools.lru_cache(maxsize=64)
def knab_window(T: float, B: float, P: int, N: int):
    """The spectral window of KnabWindow.m:11-47 for N samples at step T and bilateral bandwidth B:
    returns (G, kg), G of length 2 kg + 1 (FP64, host)."""
    df = 1.0 / (N * T)
    kg = int(math.floor((1.0 / T - B / 2) / df))
    f = -df * kg + df * np.arange(kg + 1)
    B0 = 1.0 / T
    Bw = B0 - B                       # bandwidth of the prolate window
    inside = Bw / 2 <= f + B0 / 2
    G = np.zeros(2 * kg + 1)
    G[:kg + 1][inside] = 1.0
    out = np.nonzero(~inside)[0]
    for k in range(out[-1] + 1 if out.size else 0):
        lo, hi = max(-Bw / 2, f[k] - B0 / 2), min(Bw / 2, f[k] + B0 / 2)
        if lo <= hi:
            G[k] = _conv_window_integral(lo, hi, Bw, 2 * P * T)
    G = G + G[::-1]
    G[kg] -= 1.0
    return G / N, kg


def _round(x: float) -> int:
    """MATLAB round (half away from zero)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


_plans: dict = {}
_windows: dict = {}


def _window_tables(w: tuple, device: torch.device) -> torch.Tensor:
    """The FP64 window G of knab_window(*w) as a device tensor, kept per (arguments, device)."""
    key = (w, device)
    g = _windows.get(key)
    if g is None:
        if len(_windows) >= 32:
            _windows.clear()
        g = _windows[key] = torch.from_numpy(knab_window(*w)[0]).to(device)
    return g


_MAX_PLANS = 8


def _plan(n1: int, n2: int, dtype, device: torch.device) -> Plan:
    """The type-2 plan of one (n1, n2, dtype, device), from a small least-recently-used cache: a
    plan holds an oversampled grid of up to 8192^2 x 8 B, so at most _MAX_PLANS are kept and an
    evicted plan is closed (its device memory freed) at once."""
    key = (n1, n2, dtype, device)
    p = _plans.pop(key, None)
    if p is None:
        while len(_plans) >= _MAX_PLANS:
            _plans.pop(next(iter(_plans))).close()
        p = Plan(n1, n2, dtype, device=device)
    _plans[key] = p                    # most recently used last
    return p


def clear_plan_cache() -> None:
    """Close every cached interpolation plan and window (frees their device memory)."""
    while _plans:
        _plans.pop(next(iter(_plans))).close()
    _windows.clear()
    _geoms.clear()


def knabNFFT2D(ExtrapVal, RHO_IN: torch.Tensor, COSALPHA_IN: torch.Tensor, INPUT_DATA: torch.Tensor,
               RHO_OUT: torch.Tensor, COSALPHA_OUT: torch.Tensor, B1: float, B2: float, P: int = 18) -> torch.Tensor:
    """knabNFFT2D.m:1-79 on the GPU.  RHO_IN, COSALPHA_IN: the input meshgrid (rows: cos alpha,
    columns: rho), INPUT_DATA: complex samples on it (same shape), RHO_OUT, COSALPHA_OUT: the
    output meshgrid; B1, B2: bilateral bandwidths in rho and cos alpha; returns RHO_OUT's shape,
    complex128 for complex128 samples (the library's FP64 path), complex64 otherwise.  All tensors
    on one ROCm device."""
    return _interp("knab", ExtrapVal, RHO_IN, COSALPHA_IN, INPUT_DATA, RHO_OUT, COSALPHA_OUT, B1, B2, P)


def rectNFFT2D(ExtrapVal, RHO_IN: torch.Tensor, COSALPHA_IN: torch.Tensor, INPUT_DATA: torch.Tensor,
               RHO_OUT: torch.Tensor, COSALPHA_OUT: torch.Tensor, B1: float, B2: float) -> torch.Tensor:
    """rectNFFT2D.m:1-77 on the GPU: the same steps with a rectangular spectral window (plain
    trigonometric interpolation of the samples' periodic extension), whose half-width kg comes, as
    in the reference, from the output extent in input samples rather than from the input size."""
    return _interp("rect", ExtrapVal, RHO_IN, COSALPHA_IN, INPUT_DATA, RHO_OUT, COSALPHA_OUT, B1, B2, 0)


def _interp(window: str, ExtrapVal, RHO_IN, COSALPHA_IN, INPUT_DATA, RHO_OUT, COSALPHA_OUT, B1, B2, P):
    dev = INPUT_DATA.device
    if dev.type != "cuda":
        raise ValueError(f"{window}NFFT2D: tensors must be on a ROCm device")
    if INPUT_DATA.dim() != 2 or INPUT_DATA.shape != RHO_IN.shape or RHO_IN.shape != COSALPHA_IN.shape:
        raise ValueError("The lengths of the X and Y vectors must match Z.")
    if min(INPUT_DATA.shape) < 2:
        raise ValueError("Z must be at least 2-by-2.")
    if RHO_OUT.shape != COSALPHA_OUT.shape or RHO_OUT.dim() != 2:
        raise ValueError("RHO_OUT and COSALPHA_OUT must be matrices of one shape")
    if min(RHO_OUT.shape) < 2:   # the output spacing (knabNFFT2D.m:36-37) divides by len - 1
        raise ValueError("RHO_OUT must be at least 2-by-2.")
    lib = _lib.load_library()
    st = _stream(dev)
    cdt = torch.complex128 if INPUT_DATA.dtype == torch.complex128 else torch.complex64
    ct = NFFT2D_C128 if cdt == torch.complex128 else NFFT2D_C64
    M, N = INPUT_DATA.shape
    g = _geometry(window, RHO_IN, COSALPHA_IN, RHO_OUT, COSALPHA_OUT, B1, B2, P, M, N, cdt, dev)
    a, b, len1, len2, C2 = g["a"], g["b"], g["len1"], g["len2"], g["C2"]
    data = INPUT_DATA.to(cdt).contiguous()           # M (cos alpha) x N (rho)
    # the spectrum of the samples (cos alpha x rho); the kernel below reads it transposed
    S = torch.fft.fft2(data).contiguous()
    F = torch.empty(a * N, b * M, dtype=cdt, device=dev)
    res = torch.empty(len1, len2, dtype=cdt, device=dev)
    if window == "knab":
        Gn, Gm = _window_tables(g["wn"], dev), _window_tables(g["wm"], dev)
        scale = float(N * a) * float(M * b)
    else:   # G = 1 and the samples divided by N M (rectNFFT2D.m:70)
        Gn = Gm = None
        scale = float(a) * float(b)
    ex = complex(ExtrapVal)
    with torch.cuda.device(dev):
        # repeated periodically to 2 kg + 1 frequencies, windowed (Selva eq. 25), zero padded to a N x b M
        # with frequency 0 at (a N / 2, b M / 2) (the library's centred layout) and conjugated
        check(lib.nfft2d_knab_spectrum(ct, _ptr(S), M, N, g["kgn"], g["kgm"], _ptr(Gn), _ptr(Gm), scale, _ptr(F),
                                       a * N, b * M, st), "nfft2d_knab_spectrum")
    # the tile sort only when the plan's points are not already this geometry's
    y = _plan(a * N, b * M, cdt, dev).setpts_type2_if_changed(g["x1"], g["x2"], F)
    with torch.cuda.device(dev):
        check(lib.nfft2d_knab_finish(ct, _ptr(y), len1, len2, C2, float(a * b * N * M), _ptr(g["ro"]), _ptr(g["co"]),
                                     g["strides"], g["rmin"], g["rmax"], g["cmin"], g["cmax"], ex.real, ex.imag,
                                     _ptr(res), st), "nfft2d_knab_finish")
    return res


_geoms: list = []      # geometry entries, most recently used last
_MAX_GEOMS = 2


def _geometry(window, RHO_IN, COSALPHA_IN, RHO_OUT, COSALPHA_OUT, B1, B2, P, M, N, cdt, dev) -> dict:
    """What a call computes from the grids and sizes alone: the grid steps, origins and extents
    (one reduction, one transfer to the host, which sizes the transform), the oversampling a, b,
    the window half-widths and the output points in units of the padded grid.  Kept for repeated
    calls on the same geometry: the same tensor objects, unmodified since (their version
    counters), on the same stream; such a call neither synchronises nor re-sorts the points.  An
    entry holds references to its tensors, so their memory is not reused while it stands."""
    st = _stream(dev)
    params = (window, float(B1), float(B2), int(P), M, N, cdt, dev, st.value)
    srcs = (RHO_IN, COSALPHA_IN, RHO_OUT, COSALPHA_OUT)
    vers = tuple(t._version for t in srcs)
    for i, e in enumerate(_geoms):
        if e["params"] == params and e["vers"] == vers and all(p is q for p, q in zip(e["srcs"], srcs)):
            _geoms.append(_geoms.pop(i))
            return e
    f64 = torch.float64
    lib = _lib.load_library()
    rdt = torch.float64 if cdt == torch.complex128 else torch.float32
    ct = NFFT2D_C128 if cdt == torch.complex128 else NFFT2D_C64
    rax, cax = RHO_IN[0, :], COSALPHA_IN[:, 0]
    rax, cax = (rax if rax.dtype == f64 else rax.to(f64)), (cax if cax.dtype == f64 else cax.to(f64))
    ro, co = RHO_OUT.to(f64), COSALPHA_OUT.to(f64)   # no copy for FP64 input: strided reads, meshgrid views too
    len1, len2 = ro.shape
    strides = (ctypes.c_longlong * 4)(ro.stride(0), ro.stride(1), co.stride(0), co.stride(1))
    stats = torch.empty(NFFT2D_KNAB_STATS + 4 * NFFT2D_KNAB_STAT_BLOCKS, dtype=f64, device=dev)
    with torch.cuda.device(dev):
        # the grid steps, origins, output extents and input ranges in one transfer to the host
        check(lib.nfft2d_knab_stats(_ptr(rax), rax.numel(), rax.stride(0), _ptr(cax), cax.numel(), cax.stride(0),
                                    _ptr(ro), _ptr(co), len1, len2, strides, _ptr(stats), st), "nfft2d_knab_stats")
        hs = stats.cpu().numpy()
    T1, T2, r0, c0, _, _, rmin, rmax, cmin, cmax = hs[:NFFT2D_KNAB_STATS].tolist()
    part = hs[NFFT2D_KNAB_STATS:].reshape(-1, 4)       # per block: min, max of ro; min, max of co
    nan_r, nan_c = np.isnan(part[:, :2]).any(), np.isnan(part[:, 2:]).any()
    rext = math.nan if nan_r else float(part[:, 1].max() - part[:, 0].min())
    cext = math.nan if nan_c else float(part[:, 3].max() - part[:, 2].min())
    Nn = rext / (T1 * (1 - 1 / len2))
    Mm = cext / (T2 * (1 - 1 / len1))
    if Nn < 1 or math.isnan(Nn):
        Nn = 1.0
    if Mm < 1 or math.isnan(Mm):
        Mm = 1.0
    a, b = _round(len2 / Nn), _round(len1 / Mm)
    R2, C2 = len1, len2                # extra output points at the origin when a or b < 2, cropped at the end
    if a < 2:
        a = 2
        C2 = _round(a * Nn)
    if b < 2:
        b = 2
        R2 = _round(b * Mm)
    if (a * N) % 2 or (b * M) % 2:
        raise ValueError(f"{window}NFFT2D: a*N = {a * N} and b*M = {b * M} must be even")
    if C2 < len2 or R2 < len1:
        raise ValueError(f"{window}NFFT2D: padded output grid {R2} x {C2} smaller than the output")
    wn = wm = None
    if window == "knab":
        wn, wm = (T1, float(B1), int(P), N), (T2, float(B2), int(P), M)
        kgn, kgm = knab_window(*wn)[1], knab_window(*wm)[1]
    else:   # rectNFFT2D.m:62-65: kg from the output-extent Nn, Mm
        kgn = int(math.floor((1 / T1 - float(B1) / 2) / (1 / (Nn * T1))))
        kgm = int(math.floor((1 / T2 - float(B2) / 2) / (1 / (Mm * T2))))
    if 2 * kgn + 1 > a * N or 2 * kgm + 1 > b * M or kgn < 0 or kgm < 0:
        raise ValueError(f"{window}NFFT2D: the window is wider than the padded spectrum")
    x1 = torch.empty(R2 * C2, dtype=rdt, device=dev)
    x2 = torch.empty(R2 * C2, dtype=rdt, device=dev)
    with torch.cuda.device(dev):
        # points in units of the padded grid, wrapped to one period in FP64 before the cast
        check(lib.nfft2d_knab_points(ct, _ptr(ro), _ptr(co), len1, len2, strides, R2, C2, r0, c0, T1 / a, T2 / b,
                                     float(a * N), float(b * M), _ptr(x1), _ptr(x2), st), "nfft2d_knab_points")
    e = dict(params=params, srcs=srcs, vers=vers, ro=ro, co=co, strides=strides, len1=len1, len2=len2, a=a, b=b,
             C2=C2, wn=wn, wm=wm, kgn=kgn, kgm=kgm, rmin=rmin, rmax=rmax, cmin=cmin, cmax=cmax, x1=x1, x2=x2)
    while len(_geoms) >= _MAX_GEOMS:
        _geoms.pop(0)
    _geoms.append(e)
    return e
