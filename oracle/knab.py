"""CPU restatement of the reference's band-limited polar interpolation -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s checks may import this module, and only as
the checker; the product (``miscellaneous_amd.knab``) never imports it.

It restates, step by step and in FP64 numpy:

* ``conv_window``            -- ``BP_FBP_FFBP/ConvWindow.m:5-11``;
* ``interval_intersection``  -- ``BP_FBP_FFBP/IntervalIntersection.m:4-21``;
* ``knab_window``            -- ``BP_FBP_FFBP/KnabWindow.m:11-47`` (MATLAB's
  ``quad(.., 1e-12)`` becomes ``scipy.integrate.quad`` with the same
  tolerance: both are adaptive and reach the integral to ~1e-12);
* ``selva_interp_2d``        -- ``BP_FBP_FFBP/NFFT_Based_Selvas_Approach_for_Signal_Interp_2D.m:4-28``,
  with the oracle's ``nfft1_2d`` (``NFFT1_2D.m``) as its NUFFT;
* ``knab_nfft2d``            -- ``BP_FBP_FFBP/knabNFFT2D.m:1-79``;
* ``rect_nfft2d``            -- ``BP_FBP_FFBP/rectNFFT2D.m:1-77`` (the same steps
  with a rectangular window).

Pins (``tests/test_knab.py``):

* ``conv_window`` integrates to 1 over its support (the Kaiser-Bessel pair
  int_{-1}^{1} I0(a sqrt(1-u^2)) du = 2 sinh(a)/a fixes its normalisation);
* ``knab_window`` against an independent quadrature (Gauss-Legendre after the
  substitution f = B/2 sin(theta)), and its pass band equal to 1/N;
* ``knab_nfft2d`` reproduces band-limited signals (sums of 2-D exponentials
  with frequencies inside the bands B1, B2) at off-grid points, and marks
  points outside the input grid with ExtrapVal;
* ``rect_nfft2d`` reproduces signals periodic on the input grid (frequencies on
  its DFT grid, inside the bands) at off-grid points.

Only the even-size branch of the zero padding (a N and b M even) is restated;
odd products raise, as in the GPU form.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate, special

from . import nfft1_2d


def conv_window(f, B: float, T: float):
    """Approximate prolate spheroidal window (ConvWindow.m:5-11):
    W(f) = I0(pi B T/2 sqrt(1 - (2f/B)^2)) / (B sinc(i B T/2)) for |f| < B/2, else 0,
    with MATLAB's sinc(x) = sin(pi x)/(pi x), so B sinc(i B T/2) = B sinh(pi B T/2) / (pi B T/2)."""
    f = np.asarray(f, dtype=np.float64)
    a = math.pi * B * T / 2
    w = np.zeros_like(f)
    inside = np.abs(f) < B / 2
    w[inside] = special.i0(a * np.sqrt(1 - (2 * f[inside] / B) ** 2)) / (B * math.sinh(a) / a)
    return w


def interval_intersection(ia, ib):
    """IntervalIntersection.m:4-21 (first output only): [lo, hi] or None."""
    if ia[1] < ib[0] or ib[1] < ia[0]:
        return None
    if ib[1] < ia[1]:
        return (ia[0], ib[1])
    return (ib[0], ia[1])


def knab_window(T: float, B: float, P: int, N: int):
    """KnabWindow.m:11-47: returns (G, kg), G of length 2 kg + 1."""
    df = 1.0 / (N * T)
    kg = int(math.floor((1.0 / T - B / 2) / df))
    f = -df * kg + df * np.arange(kg + 1)
    B0 = 1.0 / T
    B1 = B0 - B
    inside = B1 / 2 <= (f + B0 / 2)
    G = np.zeros(2 * kg + 1)
    G[:kg + 1][inside] = 1.0
    last = np.nonzero(~inside)[0]
    for k in range(last[-1] + 1 if last.size else 0):
        iv = interval_intersection((-B1 / 2, B1 / 2), (f[k] - B0 / 2, f[k] + B0 / 2))
        if iv is not None:
            G[k] = integrate.quad(lambda x: conv_window(np.array([x]), B1, 2 * P * T)[0], iv[0], iv[1],
                                  epsabs=1e-12, epsrel=1e-12, limit=200)[0]
    G = G + G[::-1]
    G[kg] -= 1.0
    return G / N, kg


def selva_interp_2d(data: np.ndarray, T1: float, T2: float, a: int, b: int, G: np.ndarray, kgn: int, kgm: int,
                    rho_s: np.ndarray, cos_s: np.ndarray) -> np.ndarray:
    """NFFT_Based_Selvas_Approach_for_Signal_Interp_2D.m:4-28 (data: N x M; returns the values at the
    points, in the points' order)."""
    N, M = data.shape
    if (a * N) % 2 or (b * M) % 2:
        raise NotImplementedError("odd a*N or b*M")
    sF = np.fft.fft2(data)
    sF = sF[np.mod(np.arange(-kgn, kgn + 1), N)][:, np.mod(np.arange(-kgm, kgm + 1), M)]
    sF = sF * G * (N * a) * (M * b)
    sF = np.pad(sF, ((a * N // 2 - kgn, a * N // 2 - kgn - 1), (b * M // 2 - kgm, b * M // 2 - kgm - 1)))
    x1 = np.ravel(rho_s / (T1 / a))
    x2 = np.ravel(cos_s / (T2 / b))
    return np.conj(nfft1_2d(np.conj(sF), x1, x2)) / (a * b * N * M)


def _round(x: float) -> int:
    """MATLAB round: half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def _interp_nfft2d(window: str, extrap_val, rho_in: np.ndarray, cos_in: np.ndarray, data: np.ndarray,
                   rho_out: np.ndarray, cos_out: np.ndarray, B1: float, B2: float, P: int = 18) -> np.ndarray:
    """knabNFFT2D.m:1-79 (window "knab") and rectNFFT2D.m:1-77 (window "rect"; they differ only in
    knabNFFT2D.m:62-70 / rectNFFT2D.m:62-70).  rho_in, cos_in: input meshgrid (rows: cos alpha,
    columns: rho); data: the samples on it; rho_out, cos_out: output meshgrid (len1 x len2)."""
    T1 = float(np.mean(np.diff(rho_in[0, :])))
    T2 = float(np.mean(np.diff(cos_in[:, 0])))
    rho_s = rho_out - rho_in[0, 0]
    cos_s = cos_out - cos_in[0, 0]
    len1, len2 = rho_s.shape
    N = (rho_s.max() - rho_s.min()) / (T1 * (1 - 1 / len2))
    M = (cos_s.max() - cos_s.min()) / (T2 * (1 - 1 / len1))
    if N < 1 or math.isnan(N):
        N = 1.0
    if M < 1 or math.isnan(M):
        M = 1.0
    a = _round(len2 / N)
    b = _round(len1 / M)
    if a < 2:
        a = 2
        ii = _round(a * N) - rho_s.shape[1]
        rho_s = np.pad(rho_s, ((0, 0), (0, ii)))
        cos_s = np.pad(cos_s, ((0, 0), (0, ii)))
    if b < 2:
        b = 2
        ii = _round(b * M) - rho_s.shape[0]
        rho_s = np.pad(rho_s, ((0, ii), (0, 0)))
        cos_s = np.pad(cos_s, ((0, ii), (0, 0)))
    dataT = data.T
    if window == "knab":
        Gn, kgn = knab_window(T1, B1, P, dataT.shape[0])
        Gm, kgm = knab_window(T2, B2, P, dataT.shape[1])
        G = np.outer(Gn, Gm)
    else:   # rectangular: G = 1, the samples divided by their count; kg from the output-extent N, M
        kgn = int(math.floor((1 / T1 - B1 / 2) / (1 / (N * T1))))
        kgm = int(math.floor((1 / T2 - B2 / 2) / (1 / (M * T2))))
        G = 1.0
        dataT = dataT / dataT.shape[0] / dataT.shape[1]
    res = selva_interp_2d(dataT, T1, T2, a, b, G, kgn, kgm, rho_s, cos_s).reshape(rho_s.shape)
    res = res[:len1, :len2].copy()
    xout = (rho_out[0, :] < rho_in[0, :].min()) | (rho_out[0, :] > rho_in[0, :].max())
    yout = (cos_out[:, 0] < cos_in[:, 0].min()) | (cos_out[:, 0] > cos_in[:, 0].max())
    res[:, xout] = extrap_val
    res[yout, :] = extrap_val
    return res


def knab_nfft2d(extrap_val, rho_in, cos_in, data, rho_out, cos_out, B1, B2, P: int = 18) -> np.ndarray:
    """knabNFFT2D.m:1-79: the prolate-tapered window (KnabWindow.m)."""
    return _interp_nfft2d("knab", extrap_val, rho_in, cos_in, data, rho_out, cos_out, B1, B2, P)


def rect_nfft2d(extrap_val, rho_in, cos_in, data, rho_out, cos_out, B1, B2) -> np.ndarray:
    """rectNFFT2D.m:1-77: the rectangular window (plain trigonometric interpolation)."""
    return _interp_nfft2d("rect", extrap_val, rho_in, cos_in, data, rho_out, cos_out, B1, B2)
