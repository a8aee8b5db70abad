"""The band-limited polar interpolation (BP_FBP_FFBP/knabNFFT2D.m): the oracle's
restatement pinned by mathematics, and the product's host window tables against
it.  CPU only."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import knab as oknab
from miscellaneous_amd.knab import knab_window

# scipy's adaptive quad flags the square-root edge of the window's integrand;
# the accuracy checks below hold regardless
pytestmark = pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")


def polar_like(rng, nr=96, nc=80, T1=0.05, T2=0.002):
    rho = 100 + T1 * np.arange(nc)
    cs = -0.1 + T2 * np.arange(nr)
    RHO, COS = np.meshgrid(rho, cs)
    return rho, cs, RHO, COS


def bandlimited(rng, B1, B2, q=5):
    u = (rng.random(q) - 0.5) * B1 * 0.9
    v = (rng.random(q) - 0.5) * B2 * 0.9
    c = rng.standard_normal(q) + 1j * rng.standard_normal(q)
    return lambda R, C: sum(c[k] * np.exp(2j * np.pi * (u[k] * R + v[k] * C)) for k in range(q))


@pytest.mark.parametrize("B,T", [(7.3, 0.9), (0.5, 36.0), (250.0, 0.0721)])
def test_conv_window_integrates_to_one(B, T):
    # int_{-1}^{1} I0(a sqrt(1-u^2)) du = 2 sinh(a)/a fixes ConvWindow.m's normalisation
    val = integrate.quad(lambda x: oknab.conv_window(np.array([x]), B, T)[0], -B / 2, B / 2,
                         epsabs=1e-14, epsrel=1e-14, limit=200)[0]
    assert abs(val - 1.0) < 1e-12


@pytest.mark.parametrize("T,B,P,N", [(0.05, 10.0, 18, 80), (0.002, 250.0, 18, 96), (1.0, 0.5, 18, 64),
                                     (0.3, 1.2, 10, 33)])
def test_knab_window_two_quadratures(T, B, P, N):
    """The product's Gauss-Legendre window against the oracle's adaptive quadrature."""
    G, kg = knab_window(T, B, P, N)
    Go, kgo = oknab.knab_window(T, B, P, N)
    assert kg == kgo and G.shape == (2 * kg + 1,)
    assert np.max(np.abs(G - Go)) < 1e-13 * np.max(np.abs(Go)) * N
    assert np.allclose(G, G[::-1], rtol=0, atol=1e-15)             # even window
    assert abs(G[kg] - 1.0 / N) < 1e-15                             # pass band: 1/N at f = 0


def test_oracle_reproduces_bandlimited_signals():
    """Interior off-grid points of a band-limited signal sampled at half its Nyquist rate
    (B = 0.5 / T, as FBP.m:95 calls it) are reproduced to the NFFT's FP64 accuracy."""
    rng = np.random.default_rng(0)
    rho, cs, RHO, COS = polar_like(rng)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    B1, B2 = 0.5 / T1, 0.5 / T2
    sig = bandlimited(rng, B1, B2)
    RO, CO = np.meshgrid(np.linspace(rho[20], rho[-21], 150), np.linspace(cs[20], cs[-21], 170))
    out = oknab.knab_nfft2d(0.0, RHO, COS, sig(RHO, COS), RO, CO, B1, B2)
    ref = sig(RO, CO)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-10


def test_oracle_marks_extrapolation_and_coarse_outputs():
    """Output columns / rows outside the input grid get ExtrapVal; an output grid coarser than the
    input (a, b < 2: the padded-points branch of knabNFFT2D.m:44-56) still reproduces the signal."""
    rng = np.random.default_rng(1)
    rho, cs, RHO, COS = polar_like(rng)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    B1, B2 = 0.5 / T1, 0.5 / T2
    sig = bandlimited(rng, B1, B2)
    ro = np.linspace(rho[0] - 3 * T1, rho[-1] + 2 * T1, 31)
    co = np.linspace(cs[25], cs[-26], 9)
    RO, CO = np.meshgrid(ro, co)
    out = oknab.knab_nfft2d(-7.0, RHO, COS, sig(RHO, COS), RO, CO, B1, B2)
    xout = (ro < rho.min()) | (ro > rho.max())
    assert xout.any() and np.all(out[:, xout] == -7.0)
    inner = (ro > rho[20]) & (ro < rho[-21])
    ref = sig(RO, CO)[:, inner]
    assert np.linalg.norm(out[:, inner] - ref) / np.linalg.norm(ref) < 1e-10


def test_oracle_rect_reproduces_periodic_signals():
    """rectNFFT2D.m is trigonometric interpolation of the samples' periodic extension: signals with
    frequencies on the input grid's DFT grid, inside the kept band and clear of its aliases
    (|k| < N - kg), are reproduced at off-grid points."""
    rng = np.random.default_rng(2)
    nr, nc = 96, 80
    rho, cs, RHO, COS = polar_like(rng, nr, nc)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    k1 = rng.integers(-int(0.2 * nc), int(0.2 * nc), 4)
    k2 = rng.integers(-int(0.2 * nr), int(0.2 * nr), 4)
    c = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    sig = lambda R, C: sum(c[q] * np.exp(2j * np.pi * (k1[q] * (R - rho[0]) / (nc * T1) +
                                                       k2[q] * (C - cs[0]) / (nr * T2))) for q in range(4))
    RO, CO = np.meshgrid(np.linspace(rho[0], rho[-1], 150), np.linspace(cs[0], cs[-1], 170))
    out = oknab.rect_nfft2d(0.0, RHO, COS, sig(RHO, COS), RO, CO, 0.5 / T1, 0.5 / T2)
    ref = sig(RO, CO)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-10
