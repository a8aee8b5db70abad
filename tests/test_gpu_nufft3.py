"""GPU parity of the type-3 NUFFT (NUFFT_2D_GPU, FP64) against the oracle.

Reference: NUFFT_2D_CPU, NFFT/NFFT3/C/NFFT3_2D_C.cpp:199-312.  Bars:
* the derived parameters (extents, msp, the Mry x Mrx grid, ...) bitwise equal
  to the oracle's restatement of :214-267;
* F within rel-L2 1e-11 of the oracle's step-by-step restatement (the same
  approximation; the two differ by FP64 rounding only);
* F within 10 eps of the exact FP64 sum (what the algorithm promises).
"""
import numpy as np
import pytest
import torch

import miscellaneous_amd as m
import oracle

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
TOL_RESTATEMENT = 1e-11


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def case(seed, n_in, n_out, xr, yr, sr, tr):
    rng = np.random.default_rng(seed)
    x = rng.uniform(*xr, n_in)
    y = rng.uniform(*yr, n_in)
    s = rng.uniform(*sr, n_out)
    t = rng.uniform(*tr, n_out)
    f = rng.standard_normal(n_in) + 1j * rng.standard_normal(n_in)
    return x, y, s, t, f


def run(x, y, s, t, f, eps):
    return m.NUFFT_2D(T(x), T(y), T(s), T(t), T(f), eps).cpu().numpy()


CASES = {
    "square": (11, 400, 300, (-4, 4), (-4, 4), (-5, 5), (-5, 5)),
    "offcentre": (12, 300, 350, (1, 9), (-2, -1), (-3, 12), (20, 26)),
    "anisotropic": (13, 500, 200, (-1, 1), (-6, 6), (-30, 30), (-2, 3)),
    "wide": (14, 3000, 2500, (-10, 10), (-10, 10), (-8, 8), (-8, 8)),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-12])
def test_params_bitwise(name, eps):
    x, y, s, t, _ = case(*CASES[name])
    got = m.nufft3_params(T(x), T(y), T(s), T(t), eps)
    want = oracle.nufft3_params(x, y, s, t, eps)
    print(f"case type-3 params {name} eps={eps}: Mrx={got['Mrx']} Mry={got['Mry']} msp={got['msp']}", flush=True)
    assert got == want


@pytest.mark.parametrize("name", ["square", "offcentre", "anisotropic"])
@pytest.mark.parametrize("eps", [1e-6, 1e-9, 1e-12])
def test_vs_restatement_and_direct_sum(name, eps):
    x, y, s, t, f = case(*CASES[name])
    print(f"case type-3 {name} eps={eps}", flush=True)
    F = run(x, y, s, t, f, eps)
    R, grid = oracle.nufft3_2d(x, y, s, t, f, eps)
    D = oracle.ndft3_2d(x, y, s, t, f)
    e_r, e_d = oracle.rel_l2(F, R), oracle.rel_l2(F, D)
    print(f"  grid {grid}: rel-L2 vs restatement {e_r:.2e}, vs direct sum {e_d:.2e}")
    assert e_r < TOL_RESTATEMENT
    assert e_d < 10 * eps


def test_power_of_two_grid():
    """Extents chosen so that Mrx = 256 (FFT without the chirp-z convolution) while Mry is not a power of two."""
    rng = np.random.default_rng(31)
    n = 600
    x = rng.uniform(-4, 4, n)
    x[:2] = [-4.0, 4.0]                 # X1 = 4
    s = rng.uniform(-13, 13, n)
    s[:2] = [-13.0, 13.0]               # X1 S = 52: Mrx = 2 ceil(1.5406 * 52 + 2.2 msp) = 256 at eps 1e-6
    y = rng.uniform(-3, 3, n)
    t = rng.uniform(-7, 7, n)
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    p = oracle.nufft3_params(x, y, s, t, 1e-6)
    assert p["Mrx"] == 256 and p["Mry"] & (p["Mry"] - 1)
    F = run(x, y, s, t, f, 1e-6)
    R, _ = oracle.nufft3_2d(x, y, s, t, f, 1e-6)
    assert oracle.rel_l2(F, R) < TOL_RESTATEMENT
    assert oracle.rel_l2(F, oracle.ndft3_2d(x, y, s, t, f)) < 1e-5


@pytest.mark.parametrize("eps", [1e-6, 1e-12])
def test_large_against_direct_sum_subset(eps):
    """300,000 inputs -> 300,000 outputs (grid ~1100^2 at eps 1e-6; msp = 34, two lanes per tap column at
    1e-12), 512 outputs checked against the FP64 direct sum over every input."""
    x, y, s, t, f = case(41, 300_000, 300_000, (-30, 30), (-20, 20), (-12, 12), (-15, 15))
    p = oracle.nufft3_params(x, y, s, t, eps)
    print(f"case type-3 large eps={eps}: grid {p['Mrx']} x {p['Mry']}, msp {p['msp']}", flush=True)
    F = run(x, y, s, t, f, eps)
    idx = np.random.default_rng(7).choice(len(s), 512, replace=False)
    D = oracle.ndft3_2d(x, y, s[idx], t[idx], f)
    e = oracle.rel_l2(F[idx], D)
    print(f"  rel-L2 on 512 outputs: {e:.2e}")
    assert e < 10 * eps


def test_clustered_inputs_and_outputs():
    """Most inputs in one grid bin (bins of hundreds of points: several candidate chunks per tile) and the
    outputs in two tight clusters, with spread-out anchors fixing the extents."""
    rng = np.random.default_rng(51)
    n = 5000
    x = np.concatenate([[-6.0, 6.0], 0.3 + 0.01 * rng.standard_normal(n - 2)])
    y = np.concatenate([[-5.0, 5.0], -1.1 + 0.01 * rng.standard_normal(n - 2)])
    s = np.concatenate([[-9.0, 9.0], 2.0 + 0.02 * rng.standard_normal(2000), -4 + 0.02 * rng.standard_normal(2000)])
    t = np.concatenate([[-7.0, 7.0], 0.02 * rng.standard_normal(4000)])
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    F = run(x, y, s, t, f, 1e-9)
    R, _ = oracle.nufft3_2d(x, y, s, t, f, 1e-9)
    assert oracle.rel_l2(F, R) < TOL_RESTATEMENT
    assert oracle.rel_l2(F, oracle.ndft3_2d(x, y, s, t, f)) < 1e-8


def test_host_variant_and_repeated_calls():
    """NUFFT_2D_GPU_host (host arrays, NUFFT_2D_CPU's contract) against the device entry point, and calls of
    growing then shrinking size on one shared workspace."""
    small = case(*CASES["square"])
    big = case(*CASES["wide"])
    F1 = run(*small, 1e-9)
    F2 = run(*big, 1e-9)
    F3 = run(*small, 1e-9)
    H = m.NUFFT_2D_host(*small, 1e-9)
    # inputs are summed in their arrival order inside a bin: repeated calls agree to rounding
    assert oracle.rel_l2(F1, F3) < 1e-12
    assert oracle.rel_l2(H, F1) < 1e-12
    assert oracle.rel_l2(F2, oracle.ndft3_2d(*big)) < 1e-8


def test_empty_and_single():
    x, y, s, t, f = case(*CASES["square"])
    e = np.zeros(0)
    assert m.NUFFT_2D(T(x), T(y), T(e), T(e), T(f), 1e-6).numel() == 0
    Z = m.NUFFT_2D(T(e), T(e), T(s), T(t), T(np.zeros(0, complex)), 1e-6).cpu().numpy()
    assert Z.shape == s.shape and not Z.any()
    # one weighted input between two zero-weight anchors: F = f e^{-i (s x0 + t y0)}
    xs, ys = np.array([-2.0, 1.25, 3.0]), np.array([-1.0, 0.5, 2.0])
    fs = np.array([0.0, 2 - 1j, 0.0])
    F = run(xs, ys, s, t, fs, 1e-10)
    assert oracle.rel_l2(F, (2 - 1j) * np.exp(-1j * (s * 1.25 + t * 0.5))) < 1e-9


def test_errors():
    x, y, s, t, f = case(*CASES["square"])
    for eps in (0.0, 1.0, -1e-6, float("nan")):
        with pytest.raises(m.NFFT2DError, match="eps"):
            m.NUFFT_2D(T(x), T(y), T(s), T(t), T(f), eps)
    with pytest.raises(m.NFFT2DError, match="extent"):
        m.NUFFT_2D(T(np.full_like(x, 0.5)), T(y), T(s), T(t), T(f), 1e-6)
    bad = s.copy()
    bad[3] = np.nan
    with pytest.raises(m.NFFT2DError, match="finite"):
        m.NUFFT_2D(T(x), T(y), T(bad), T(t), T(f), 1e-6)
    with pytest.raises(m.NFFT2DError, match="taps"):
        m.NUFFT_2D(T(x), T(y), T(s), T(t), T(f), 1e-16)
    with pytest.raises(m.NFFT2DError, match="grid"):
        m.NUFFT_2D(T(x * 100), T(y), T(s * 100), T(t), T(f), 1e-6)
    # still usable after the errors
    assert oracle.rel_l2(run(x, y, s, t, f, 1e-9), oracle.ndft3_2d(x, y, s, t, f)) < 1e-8
