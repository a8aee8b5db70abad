"""GPU parity of knabNFFT2D (BP_FBP_FFBP/knabNFFT2D.m) against the oracle's
FP64 restatement and against band-limited signals.  Tolerance: rel-L2 1e-5
(the FP32 path)."""
import numpy as np
import pytest
import torch

import miscellaneous_amd as m
from oracle import knab as oknab

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")]

DEV = torch.device("cuda:0")
TOL = 1e-5


def T(a, dt=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t if dt is None else t.to(dt)


def grid(nr, nc, T1=0.05, T2=0.002):
    rho = 100 + T1 * np.arange(nc)
    cs = -0.1 + T2 * np.arange(nr)
    RHO, COS = np.meshgrid(rho, cs)
    return rho, cs, RHO, COS


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nr,nc,lo,hi", [(96, 80, 150, 170), (64, 128, 300, 90), (50, 60, 40, 45)])
def test_knab_matches_oracle(nr, nc, lo, hi):
    """Random (not band-limited) samples, denser and coarser output grids, some output
    points outside the input grid (ExtrapVal)."""
    rng = np.random.default_rng(nr + nc)
    rho, cs, RHO, COS = grid(nr, nc)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    data = ((rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))) / np.sqrt(2))
    RO, CO = np.meshgrid(np.linspace(rho[0] - 2 * T1, rho[-1], lo), np.linspace(cs[3], cs[-1] + T2, hi))
    print(f"case knab {nr}x{nc} -> {hi}x{lo}", flush=True)
    ref = oknab.knab_nfft2d(0.5, RHO, COS, data, RO, CO, 0.5 / T1, 0.5 / T2)
    out = m.knabNFFT2D(0.5, T(RHO), T(COS), T(data, torch.complex64), T(RO), T(CO), 0.5 / T1, 0.5 / T2)
    out = out.cpu().numpy()
    assert out.shape == ref.shape
    inside = ref != 0.5
    assert np.all(out[~inside] == 0.5)
    e = rel(out[inside], ref[inside])
    print(f"knab rel-L2 {e:.3e}")
    assert e < TOL


def test_knab_reproduces_bandlimited_signal():
    rng = np.random.default_rng(5)
    rho, cs, RHO, COS = grid(128, 128)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    B1, B2 = 0.5 / T1, 0.5 / T2
    u = (rng.random(6) - 0.5) * B1 * 0.9
    v = (rng.random(6) - 0.5) * B2 * 0.9
    c = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    sig = lambda R, C: sum(c[k] * np.exp(2j * np.pi * (u[k] * R + v[k] * C)) for k in range(6))
    RO, CO = np.meshgrid(np.linspace(rho[20], rho[-21], 257), np.linspace(cs[20], cs[-21], 255))
    out = m.knabNFFT2D(0.0, T(RHO), T(COS), T(sig(RHO, COS), torch.complex64), T(RO), T(CO), B1, B2).cpu().numpy()
    assert rel(out, sig(RO, CO)) < TOL


def test_knab_rejects_odd_padded_sizes():
    rho, cs, RHO, COS = grid(51, 61)                 # odd N with odd a
    data = np.ones((51, 61), dtype=np.complex64)
    RO, CO = np.meshgrid(np.linspace(rho[0], rho[-1], 183), np.linspace(cs[0], cs[-1], 102))
    with pytest.raises(ValueError, match="must be even"):
        m.knabNFFT2D(0.0, T(RHO), T(COS), T(data), T(RO), T(CO), 10.0, 250.0)


@pytest.mark.parametrize("nr,nc,lo,hi", [(96, 80, 150, 170), (50, 60, 40, 45)])
def test_rect_matches_oracle(nr, nc, lo, hi):
    """rectNFFT2D (rectangular window) against the oracle, dense and coarse outputs."""
    rng = np.random.default_rng(nr * nc)
    rho, cs, RHO, COS = grid(nr, nc)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    data = ((rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))) / np.sqrt(2))
    RO, CO = np.meshgrid(np.linspace(rho[0] - 2 * T1, rho[-1], lo), np.linspace(cs[3], cs[-1] + T2, hi))
    print(f"case rect {nr}x{nc} -> {hi}x{lo}", flush=True)
    ref = oknab.rect_nfft2d(0.5, RHO, COS, data, RO, CO, 0.5 / T1, 0.5 / T2)
    out = m.rectNFFT2D(0.5, T(RHO), T(COS), T(data, torch.complex64), T(RO), T(CO), 0.5 / T1, 0.5 / T2)
    out = out.cpu().numpy()
    inside = ref != 0.5
    assert np.all(out[~inside] == 0.5)
    e = rel(out[inside], ref[inside])
    print(f"rect rel-L2 {e:.3e}")
    assert e < TOL


@pytest.mark.parametrize("fn", ["knab", "rect"])
def test_fp64_samples_take_the_fp64_path(fn):
    """complex128 samples run on complex128 plans: the oracle's FP64 values to 1e-9."""
    rng = np.random.default_rng(21)
    nr, nc = 96, 80
    rho, cs, RHO, COS = grid(nr, nc)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    data = ((rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))) / np.sqrt(2))
    RO, CO = np.meshgrid(np.linspace(rho[2], rho[-3], 150), np.linspace(cs[3], cs[-4], 170))
    ref = getattr(oknab, f"{fn}_nfft2d")(0.0, RHO, COS, data, RO, CO, 0.5 / T1, 0.5 / T2)
    out = getattr(m, f"{fn}NFFT2D")(0.0, T(RHO), T(COS), T(data), T(RO), T(CO), 0.5 / T1, 0.5 / T2)
    assert out.dtype == torch.complex128
    e = rel(out.cpu().numpy(), ref)
    print(f"{fn} fp64 rel-L2 {e:.3e}")
    assert e < 1e-9


@pytest.mark.parametrize("fn", ["knab", "rect"])
def test_meshgrid_views_and_fp32_coordinates(fn):
    """Output coordinates as torch.meshgrid views (stride 0 in one dimension) and in FP32: the
    kernels read them in place; the result is bitwise that of contiguous FP64 copies of the same
    values."""
    n1, n2 = 80, 96
    T1, T2 = 0.05, 0.002
    f64 = torch.float64
    rho = 100 + T1 * torch.arange(n2, dtype=f64, device=DEV)
    cs = -0.1 + T2 * torch.arange(n1, dtype=f64, device=DEV)
    RHO, COS = torch.meshgrid(rho, cs, indexing="xy")          # views, n1 x n2
    assert RHO.stride(0) == 0 or COS.stride(1) == 0
    rng = np.random.default_rng(33)
    data = T((rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))) / np.sqrt(2), torch.complex64)
    ro = torch.linspace(float(rho[0]) - 2 * T1, float(rho[-1]), 150, device=DEV).to(torch.float32)
    co = torch.linspace(float(cs[3]), float(cs[-1]) + T2, 170, device=DEV).to(torch.float32)
    RO, CO = torch.meshgrid(ro, co, indexing="xy")
    f = getattr(m, f"{fn}NFFT2D")
    a = f(0.25, RHO, COS, data, RO, CO, 0.5 / T1, 0.5 / T2)
    b = f(0.25, RHO.contiguous(), COS.contiguous(), data, RO.to(f64).contiguous(), CO.to(f64).contiguous(),
          0.5 / T1, 0.5 / T2)
    assert torch.equal(a, b)
    assert (a == 0.25).any() and (a != 0.25).any()


def test_repeated_geometry_reuses_points_and_sees_in_place_changes():
    """Repeated calls on the same grid tensors reuse the grid statistics, the output points and
    the plan's tile sort: a second call equals the first bitwise, new samples on the same
    geometry match the oracle, an in-place change of the output grid is seen (version counter),
    and a same-shaped output grid in new tensors is computed afresh."""
    rng = np.random.default_rng(77)
    nr, nc = 64, 96
    rho, cs, RHO, COS = grid(nr, nc)
    T1, T2 = rho[1] - rho[0], cs[1] - cs[0]
    B1, B2 = 0.5 / T1, 0.5 / T2
    d1 = (rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))) / np.sqrt(2)
    d2 = (rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))) / np.sqrt(2)
    RO, CO = np.meshgrid(np.linspace(rho[2], rho[-3], 120), np.linspace(cs[2], cs[-3], 90))
    tR, tC, tRO, tCO = T(RHO), T(COS), T(RO), T(CO)
    print("case knab repeated geometry 64x96 -> 90x120", flush=True)
    a = m.knabNFFT2D(0.0, tR, tC, T(d1, torch.complex64), tRO, tCO, B1, B2)
    b = m.knabNFFT2D(0.0, tR, tC, T(d1, torch.complex64), tRO, tCO, B1, B2)
    assert torch.equal(a, b)
    c = m.knabNFFT2D(0.0, tR, tC, T(d2, torch.complex64), tRO, tCO, B1, B2).cpu().numpy()
    assert rel(c, oknab.knab_nfft2d(0.0, RHO, COS, d2, RO, CO, B1, B2)) < TOL
    # the same tensors changed in place: new points
    RO2, CO2 = np.meshgrid(np.linspace(rho[5], rho[-8], 120), np.linspace(cs[4], cs[-6], 90))
    tRO.copy_(T(RO2))
    tCO.copy_(T(CO2))
    e = m.knabNFFT2D(0.0, tR, tC, T(d2, torch.complex64), tRO, tCO, B1, B2).cpu().numpy()
    assert rel(e, oknab.knab_nfft2d(0.0, RHO, COS, d2, RO2, CO2, B1, B2)) < TOL
    # new tensors of the first grid: the first geometry's result again
    f = m.knabNFFT2D(0.0, T(RHO), T(COS), T(d1, torch.complex64), T(RO), T(CO), B1, B2)
    assert rel(f.cpu().numpy(), a.cpu().numpy()) < 1e-6
