"""Pins for the CPU oracle (the checker every GPU result is compared with).

Each function of the oracle is checked against something it was not derived
from: 50-digit mpmath closed forms, scipy/numpy library routines, committed
FP64 direct-sum golden vectors, invariants of the transform (adjoint,
shift theorem, periodicity) and the reference's own known-answer relation
NFFT/NFFT2/Matlab/Test_NFFT2_2D.m:24-29.  Planted mistakes in the oracle
(``mutation``) must fail these checks.
"""
import numpy as np
import pytest
import scipy.integrate
import scipy.special

import oracle

K = 6
C = 2
TOL_FAST = 1e-10      # c=2, K=6 has an intrinsic error of ~5e-12 rel-L2


# ------------------------------------------------------------------ scalar pieces
def test_alpha_matches_reference_constant(golden):
    w = golden("window_mp.npz")
    # NFFT/NFFT2_2D.cu:27: (2 - 1/cc) * pi - 0.01
    assert oracle.alpha(2) == pytest.approx(float(w["alpha"]), rel=1e-15)
    assert oracle.alpha(2) == pytest.approx(1.5 * np.pi - 0.01, rel=1e-15)


def test_bessel_i0_vs_mpmath_and_scipy(golden):
    w = golden("window_mp.npz")
    for x, ref in zip(w["i0_x"], w["i0"]):
        assert oracle.bessel_i0(x) == pytest.approx(ref, rel=2e-15)
        assert oracle.bessel_i0(x) == pytest.approx(float(scipy.special.i0(x)), rel=1e-13)


def test_spectral_window_vs_mpmath(golden):
    w = golden("window_mp.npz")
    for N in (16, 22, 128):
        np.testing.assert_allclose(oracle.spectral_window(N), w[f"spec_{N}"], rtol=1e-14)


def test_window_vs_mpmath(golden):
    w = golden("window_mp.npz")
    got = oracle.window(w["d"])
    np.testing.assert_allclose(got, w["phi"], rtol=1e-13, atol=1e-12 * np.abs(w["phi"]).max())
    got_s = oracle.window(w["d"], series=True)
    np.testing.assert_allclose(got_s, w["phi"], rtol=1e-13, atol=1e-12 * np.abs(w["phi"]).max())


def test_window_is_fourier_pair_of_spectral_window():
    """phi_hat(d) = (1/2pi) int_{-alpha}^{alpha} I0(K sqrt(alpha^2 - xi^2)) cos(xi d) dxi."""
    a = oracle.alpha(C)
    for d in (0.0, 0.37, 1.3, 4.0, 5.9, 6.0, 6.3):
        val, _ = scipy.integrate.quad(lambda xi: scipy.special.i0(K * np.sqrt(max(a * a - xi * xi, 0.0)))
                                      * np.cos(xi * d), -a, a, limit=400, epsabs=0, epsrel=1e-12)
        ref = val / (2 * np.pi)
        scale = oracle.window([0.0])[0]
        assert abs(oracle.window([d])[0] - ref) <= 1e-9 * scale


def test_window_deriv_vs_mpmath(golden):
    w = golden("window_mp.npz")
    got = oracle.window_deriv(w["d"])
    np.testing.assert_allclose(got, w["dphi"], rtol=1e-9, atol=1e-10 * np.abs(w["dphi"]).max())


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (44, 44), (12, 20), (6, 10)])
def test_fft2_vs_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    np.testing.assert_allclose(oracle.fft2(a), np.fft.fft2(a), rtol=0, atol=1e-12 * np.abs(a).sum())
    np.testing.assert_allclose(oracle.fft2(a, sign=+1), np.fft.ifft2(a) * a.size, rtol=0,
                               atol=1e-12 * np.abs(a).sum())


def test_fft2_impulse_and_exponential():
    L1, L2 = 16, 8
    a = np.zeros((L1, L2), complex)
    a[3, 5] = 1.0
    n1, n2 = np.meshgrid(np.arange(L1), np.arange(L2), indexing="ij")
    np.testing.assert_allclose(oracle.fft2(a), np.exp(-2j * np.pi * (3 * n1 / L1 + 5 * n2 / L2)), atol=1e-14)
    e = np.exp(2j * np.pi * (2 * n1 / L1 + 1 * n2 / L2))
    ref = np.zeros((L1, L2), complex)
    ref[2, 1] = L1 * L2
    np.testing.assert_allclose(oracle.fft2(e), ref, atol=1e-12)


# ------------------------------------------------------------------ full transforms
@pytest.mark.parametrize("tag", ["a", "b", "c"])
@pytest.mark.parametrize("round_mode", [0, 1])
def test_type2_vs_golden_direct_sum(golden, tag, round_mode):
    g = golden("tiny.npz")
    f = g[f"{tag}_f"].astype(np.complex128)
    y = oracle.nfft1_2d(f, g[f"{tag}_x1"], g[f"{tag}_x2"], round_mode=round_mode)
    assert oracle.rel_l2(y, g[f"{tag}_y"]) < TOL_FAST


@pytest.mark.parametrize("tag", ["a", "b", "c"])
@pytest.mark.parametrize("round_mode", [0, 1])
def test_type1_vs_golden_direct_sum(golden, tag, round_mode):
    g = golden("tiny.npz")
    f = g[f"{tag}_f"]
    N1, N2 = f.shape
    F = oracle.nfft2_2d(g[f"{tag}_d"].astype(np.complex128), g[f"{tag}_x1"], g[f"{tag}_x2"], N1, N2,
                        round_mode=round_mode)
    assert oracle.rel_l2(F, g[f"{tag}_F"]) < TOL_FAST


def test_c1_type2_and_type1_vs_golden(golden):
    g = golden("c1_128_16384.npz")
    y = oracle.nfft1_2d(g["f"].astype(np.complex128), g["x1"], g["x2"])
    assert oracle.rel_l2(y, g["y"]) < TOL_FAST
    F = oracle.nfft2_2d(g["d"].astype(np.complex128), g["x1"], g["x2"], 128, 128)
    assert oracle.rel_l2(F, g["F"]) < TOL_FAST


def test_fp32_variant_within_single_precision(golden):
    g = golden("c1_128_16384.npz")
    y = oracle.nfft1_2d(g["f"], g["x1"], g["x2"], precision="float")
    assert oracle.rel_l2(y, g["y"]) < 2e-6
    F = oracle.nfft2_2d(g["d"], g["x1"], g["x2"], 128, 128, precision="float")
    assert oracle.rel_l2(F, g["F"]) < 2e-6


def test_direct_sums_vs_golden(golden):
    g = golden("tiny.npz")
    for tag in "abc":
        f = g[f"{tag}_f"].astype(np.complex128)
        y, d1, d2 = oracle.ndft_type2(f, g[f"{tag}_x1"], g[f"{tag}_x2"], with_dx=True)
        assert oracle.rel_l2(y, g[f"{tag}_y"]) < 1e-12
        assert oracle.rel_l2(d1, g[f"{tag}_dy1"]) < 1e-12
        assert oracle.rel_l2(d2, g[f"{tag}_dy2"]) < 1e-12
        N1, N2 = f.shape
        F = oracle.ndft_type1(g[f"{tag}_d"].astype(np.complex128), g[f"{tag}_x1"], g[f"{tag}_x2"], N1, N2)
        assert oracle.rel_l2(F, g[f"{tag}_F"]) < 1e-12


def test_type1_modes_vs_golden(golden):
    """ndft_type1_modes (the FP64 checker of the GPU type-1 at selected modes) against the committed
    golden direct sums F (tests/golden/make_golden.py, numpy), at seeded mode indices, on square and
    non-square cases: a swapped (i1, i2) or (N1, N2) reads another mode and fails."""
    g = golden("tiny.npz")
    c1 = golden("c1_128_16384.npz")
    cases = [(g[f"{t}_d"], g[f"{t}_x1"], g[f"{t}_x2"], g[f"{t}_F"]) for t in "abc"]
    cases.append((c1["d"], c1["x1"], c1["x2"], c1["F"]))
    rng = np.random.default_rng(11)
    for d, x1, x2, F in cases:
        N1, N2 = F.shape
        q = 256
        i1 = rng.integers(0, N1, q)
        i2 = rng.integers(0, N2, q)
        got = oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i1, i2)
        assert oracle.rel_l2(got, F[i1, i2]) < 1e-12
        if N1 != N2:
            # plausible slips fail: the two mode indices swapped, or the two periods swapped
            sw = oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i2 % N1, i1 % N2)
            assert oracle.rel_l2(sw, F[i1, i2]) > 1e-2
            sn = oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N2, N1, i1 % N2, i2 % N1)
            assert oracle.rel_l2(sn, F[i1 % N2, i2 % N1]) > 1e-2
    # and against the full direct sum on a ragged set (one point, an empty mode list)
    d, x1, x2, F = cases[1]
    N1, N2 = F.shape
    full = oracle.ndft_type1(d[:1].astype(np.complex128), x1[:1], x2[:1], N1, N2)
    ii1, ii2 = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    one = oracle.ndft_type1_modes(d[:1].astype(np.complex128), x1[:1], x2[:1], N1, N2, ii1.ravel(), ii2.ravel())
    assert oracle.rel_l2(one, full.ravel()) < 1e-14
    assert oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, [], []).shape == (0,)


def test_derivative_vs_central_differences():
    rng = np.random.default_rng(7)
    N1, N2, M = 16, 16, 64
    f = rng.standard_normal((N1, N2)) + 1j * rng.standard_normal((N1, N2))
    x1 = rng.uniform(-8, 8, M)
    x2 = rng.uniform(-8, 8, M)
    _, d1, d2 = oracle.ndft_type2(f, x1, x2, with_dx=True)
    h = 1e-3
    fd1 = (oracle.ndft_type2(f, x1 + h, x2) - oracle.ndft_type2(f, x1 - h, x2)) / (2 * h)
    fd2 = (oracle.ndft_type2(f, x1, x2 + h) - oracle.ndft_type2(f, x1, x2 - h)) / (2 * h)
    assert oracle.rel_l2(d1, fd1) < 1e-4     # O(h^2 (2 pi k/N)^2 / 6)
    assert oracle.rel_l2(d2, fd2) < 1e-4


def test_ndft2_2d_vs_golden(golden):
    g = golden("ndft.npz")
    for p in ("g", "m"):
        out = oracle.ndft2_2d(g[f"{p}_X"], g[f"{p}_Y"], g[f"{p}_u"], g[f"{p}_v"], g[f"{p}_in"],
                              int(g[f"{p}_Nu"]), int(g[f"{p}_Nv"]))
        assert oracle.rel_l2(out, g[f"{p}_out"]) < 1e-13


def test_reference_known_answer_relation(golden):
    """Test_NFFT2_2D.m:24-29: NDFT2_2D(data,x,y,N,M) vs NFFT2_2D(data,y,x,M,N)."""
    g = golden("kat_test_nfft2_2d.npz")
    N, M = int(g["N"]), int(g["M"])
    t2 = oracle.nfft2_2d(g["data"].astype(np.complex128), g["y"], g["x"], M, N, round_mode=1)
    assert 100 * oracle.rel_l2(t2, g["transf1"]) < 1e-8      # the script prints 100*rel-L2
    t2r = oracle.nfft2_2d(g["data"].astype(np.complex128), g["y"], g["x"], M, N, round_mode=0)
    assert 100 * oracle.rel_l2(t2r, g["transf1"]) < 1e-8


# ------------------------------------------------------------------ invariants
def _case(seed=3, N1=16, N2=24, M=200):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N1, N2)) + 1j * rng.standard_normal((N1, N2))
    x1 = rng.uniform(-N1 / 2, N1 / 2, M)
    x2 = rng.uniform(-N2 / 2, N2 / 2, M)
    d = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return f, x1, x2, d


def test_adjoint_identity():
    """<A f, g> = <f, A^H g> with A^H g = conj(type1(conj g)) (the reference's
    own conj trick, BP_FBP_FFBP/NFFT_Based_Selvas_Approach_for_Signal_Interp_2D.m:28)."""
    f, x1, x2, g = _case()
    N1, N2 = f.shape
    Af = oracle.nfft1_2d(f, x1, x2)
    AHg = np.conj(oracle.nfft2_2d(np.conj(g), x1, x2, N1, N2))
    lhs = np.vdot(g, Af)
    rhs = np.vdot(AHg, f)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs) * 10


def test_shift_theorem_and_periodicity():
    f, x1, x2, d = _case()
    N1, N2 = f.shape
    k1 = np.arange(-N1 // 2, N1 // 2)
    y = oracle.nfft1_2d(f, x1, x2)
    y_shift = oracle.nfft1_2d(f, x1 + 1.0, x2)
    y_mod = oracle.nfft1_2d(f * np.exp(-2j * np.pi * k1 / N1)[:, None], x1, x2)
    assert oracle.rel_l2(y_shift, y_mod) < TOL_FAST
    assert oracle.rel_l2(oracle.nfft1_2d(f, x1 + N1, x2 - 2 * N2), y) < TOL_FAST
    F = oracle.nfft2_2d(d, x1, x2, N1, N2)
    assert oracle.rel_l2(oracle.nfft2_2d(d, x1 - N1, x2 + N2, N1, N2), F) < TOL_FAST


def test_single_point_and_single_mode_closed_forms():
    N1, N2 = 16, 12
    k1 = np.arange(-N1 // 2, N1 // 2)[:, None]
    k2 = np.arange(-N2 // 2, N2 // 2)[None, :]
    x, y = 3.37, -4.81
    F = oracle.nfft2_2d(np.array([1.0 + 0j]), np.array([x]), np.array([y]), N1, N2)
    ref = np.exp(-2j * np.pi * (k1 * x / N1 + k2 * y / N2))
    assert oracle.rel_l2(F, ref) < TOL_FAST
    f = np.zeros((N1, N2), complex)
    f[N1 // 2 + 3, N2 // 2 - 5] = 1.0
    rng = np.random.default_rng(5)
    x1 = rng.uniform(-8, 8, 50)
    x2 = rng.uniform(-6, 6, 50)
    yy = oracle.nfft1_2d(f, x1, x2)
    assert oracle.rel_l2(yy, np.exp(-2j * np.pi * (3 * x1 / N1 - 5 * x2 / N2))) < TOL_FAST


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_mistakes_are_caught(golden, mutation):
    g = golden("tiny.npz")
    f = g["b_f"].astype(np.complex128)
    N1, N2 = f.shape
    y = oracle.nfft1_2d(f, g["b_x1"], g["b_x2"], mutation=mutation)
    assert oracle.rel_l2(y, g["b_y"]) > 1e-2
    F = oracle.nfft2_2d(g["b_d"].astype(np.complex128), g["b_x1"], g["b_x2"], N1, N2, mutation=mutation)
    assert oracle.rel_l2(F, g["b_F"]) > 1e-2


# ------------------------------------------------------------------ 1-D transforms
def _numpy_ndft1d(N, x, f=None, d=None):
    """The BLAS forms NFFT1_1D_BLAS.m / NFFT2_1D_BLAS.m in grid units, as numpy matrix products."""
    k = np.arange(-N // 2, N // 2)
    E = np.exp(-2j * np.pi * np.outer(x, k) / N)       # [l][k]
    return E @ f if f is not None else E.T @ d


@pytest.mark.parametrize("N,M", [(16, 40), (64, 300), (22, 57)])
def test_1d_direct_sums_vs_numpy(N, M):
    rng = np.random.default_rng(N + M)
    x = rng.uniform(-N, N, M)
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    d = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    assert oracle.rel_l2(oracle.ndft1d_type2(f, x), _numpy_ndft1d(N, x, f=f)) < 1e-13
    assert oracle.rel_l2(oracle.ndft1d_type1(d, x, N), _numpy_ndft1d(N, x, d=d)) < 1e-13


@pytest.mark.parametrize("N,M", [(16, 40), (64, 300), (22, 57), (128, 1000)])
@pytest.mark.parametrize("round_mode", [0, 1])
def test_1d_restatements_vs_direct_sums(N, M, round_mode):
    """NFFT1_1D.m / NFFT2_1D.m restated against their direct sums (ties and points a period out
    included: x is drawn over two periods and half the points sit on half-integer c x)."""
    rng = np.random.default_rng(3 * N + M)
    x = rng.uniform(-N, N, M)
    x[: M // 4] = np.round(2 * x[: M // 4] + 0.5) / 2 - 0.25          # c x on a .5 tie
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    d = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    assert oracle.rel_l2(oracle.nfft1_1d(f, x, round_mode=round_mode), oracle.ndft1d_type2(f, x)) < TOL_FAST
    assert oracle.rel_l2(oracle.nfft2_1d(d, x, N, round_mode=round_mode), oracle.ndft1d_type1(d, x, N)) < TOL_FAST


def test_1d_adjoint_and_2d_consistency():
    """<A f, g> = <f, conj(type1(conj g))> between the two 1-D restatements, and the 1-D type-2 equals
    the 2-D restatement on a single mode row (k1 = 0 of a 2 x N array, x1 = 0)."""
    rng = np.random.default_rng(77)
    N, M = 32, 200
    x = rng.uniform(-N / 2, N / 2, M)
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    g = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    lhs = np.vdot(g, oracle.nfft1_1d(f, x))
    rhs = np.vdot(np.conj(oracle.nfft2_1d(np.conj(g), x, N)), f)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    f2 = np.zeros((2, N), complex)
    f2[1] = f                                              # k1 = 0 is row 1 of 2
    y2 = oracle.nfft1_2d(f2, np.zeros(M), x, round_mode=1)
    assert oracle.rel_l2(oracle.nfft1_1d(f, x), y2) < TOL_FAST



# ---------------------------------------------------------------- type-3 (NFFT3_2D_C.cpp)
def _type3_case(seed, n_in, n_out, xr, yr, sr, tr):
    rng = np.random.default_rng(seed)
    x = rng.uniform(*xr, n_in)
    y = rng.uniform(*yr, n_in)
    s = rng.uniform(*sr, n_out)
    t = rng.uniform(*tr, n_out)
    f = rng.standard_normal(n_in) + 1j * rng.standard_normal(n_in)
    return x, y, s, t, f


def test_ndft3_vs_numpy():
    x, y, s, t, f = _type3_case(5, 60, 45, (-3, 2), (0, 4), (-7, 9), (-2, 2))
    ref = np.exp(-1j * (np.outer(s, x) + np.outer(t, y))) @ f
    assert oracle.rel_l2(oracle.ndft3_2d(x, y, s, t, f), ref) < 1e-13


@pytest.mark.parametrize("eps", [1e-6, 1e-9, 1e-12])
@pytest.mark.parametrize("case", [(11, 400, 300, (-4, 4), (-4, 4), (-5, 5), (-5, 5)),
                                  (12, 300, 350, (1, 9), (-2, -1), (-3, 12), (20, 26)),   # off-centre, anisotropic
                                  (13, 500, 200, (-1, 1), (-6, 6), (-30, 30), (-2, 3))])
def test_nufft3_restatement_vs_direct_sum(eps, case):
    """NUFFT_2D_CPU restated step by step approximates the exact type-3 sum to ~eps (relative L2;
    the bound is 10 eps, the measured errors are 0.01-1 eps)."""
    x, y, s, t, f = _type3_case(*case)
    F, (Mrx, Mry) = oracle.nufft3_2d(x, y, s, t, f, eps)
    assert oracle.rel_l2(F, oracle.ndft3_2d(x, y, s, t, f)) < 10 * eps
    assert Mrx % 2 == 0 and Mry % 2 == 0


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_nufft3_planted_mistakes_are_caught(mutation):
    x, y, s, t, f = _type3_case(11, 400, 300, (-4, 4), (-4, 4), (-5, 5), (-5, 5))
    F, _ = oracle.nufft3_2d(x, y, s, t, f, 1e-9, mutation=mutation)
    assert oracle.rel_l2(F, oracle.ndft3_2d(x, y, s, t, f)) > 1e-2


def test_nufft3_single_input_closed_form():
    """One input f at (x0, y0) among two anchors of zero weight: F_j = f e^{-i (s_j x0 + t_j y0)}."""
    rng = np.random.default_rng(21)
    s, t = rng.uniform(-6, 6, 100), rng.uniform(-3, 8, 100)
    x = np.array([-2.0, 1.25, 3.0])
    y = np.array([-1.0, 0.5, 2.0])
    f = np.array([0.0, 2 - 1j, 0.0])
    F, _ = oracle.nufft3_2d(x, y, s, t, f, 1e-10)
    exact = (2 - 1j) * np.exp(-1j * (s * 1.25 + t * 0.5))
    assert oracle.rel_l2(F, exact) < 1e-9


def test_nufft3_params_invariants():
    x, y, s, t, f = _type3_case(12, 300, 350, (1, 9), (-2, -1), (-3, 12), (20, 26))
    for eps in (1e-3, 1e-6, 1e-12):
        p = oracle.nufft3_params(x, y, s, t, eps)
        X1, Y1 = (x.max() - x.min()) / 2, (y.max() - y.min()) / 2
        S, T = (s.max() - s.min()) / 2, (t.max() - t.min()) / 2
        msp_raw = p["msp"] + 2                                # msp = floor(msp_raw - 2)
        assert p["Mrx"] % 2 == 0 and p["Mry"] % 2 == 0
        assert p["Mrx"] >= 2 * (X1 * S / np.pi * 2.2 ** 2 + 2.2 * msp_raw) - 1e-9
        assert p["Mry"] >= 2 * (Y1 * T / np.pi * 2.2 ** 2 + 2.2 * msp_raw) - 1e-9
        # x / kk_x spans pi (half-extent pi / 2), s * kk_x spans 2 S kk_x
        assert abs((x.max() - x.min()) / p["kkx"] - np.pi) < 1e-12
        assert abs(p["Dx"] * p["Mrx"] - 2 * np.pi) < 1e-12
        assert abs(p["t1"] * 4 * p["b"] - 1) < 1e-14
