"""GPU parity of the batched 1-D NFFTs (nfft1d_*) against the oracle.

Type-2 is BP_FBP_FFBP/NFFT1_1D.m, type-1 NFFT/NFFT2/1D/Matlab/NFFT2_1D.m, both
complex64 (FP32 arithmetic): rel-L2 1e-5 per line against the FP64 direct sums
and the FP64 restatements of those files.  Coordinates are fp32 values given
to the oracle unchanged.
"""
import numpy as np
import pytest
import torch

import miscellaneous_amd as m
import oracle

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
TOL = 1e-5


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def lines(rng, B, N, M, spread=1.0):
    x = (rng.uniform(-spread, spread, (B, M)) * N).astype(np.float32)
    x[:, :4] = np.array([0.25, -0.75, N / 2 - 0.25, -N / 2], np.float32)   # c x on ties, both ends of the period
    f = ((rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))) / np.sqrt(2)).astype(np.complex64)
    d = ((rng.standard_normal((B, M)) + 1j * rng.standard_normal((B, M))) / np.sqrt(2)).astype(np.complex64)
    return x, f, d


@pytest.mark.parametrize("N,M,B", [(64, 500, 7), (512, 3000, 3), (4096, 5000, 2), (40, 300, 5), (88, 1936, 4),
                                   (8, 50, 3), (2046, 2000, 2)])
def test_type2_lines(N, M, B):
    rng = np.random.default_rng(N * 10 + M + B)
    x, f, _ = lines(rng, B, N, M)
    print(f"case 1-D type-2 N={N} M={M} lines={B}", flush=True)
    y = m.Plan1D(N, device=DEV).type2(T(f), T(x)).cpu().numpy()
    for b in range(B):
        e = oracle.rel_l2(y[b], oracle.ndft1d_type2(f[b].astype(np.complex128), x[b]))
        assert e < TOL, (b, e)
        assert oracle.rel_l2(y[b], oracle.nfft1_1d(f[b].astype(np.complex128), x[b], round_mode=0)) < TOL


@pytest.mark.parametrize("N,M,B", [(64, 500, 7), (512, 3000, 3), (2048, 5000, 2), (40, 300, 5), (88, 1936, 4),
                                   (8, 50, 3), (2046, 2000, 2)])
def test_type1_lines(N, M, B):
    rng = np.random.default_rng(N * 11 + M + B)
    x, _, d = lines(rng, B, N, M, spread=0.75)
    print(f"case 1-D type-1 N={N} M={M} lines={B}", flush=True)
    plan = m.Plan1D(N, device=DEV)
    F = plan.type1(T(d), T(x))
    F2 = plan.type1(T(d), T(x))
    assert torch.equal(F, F2)                     # fixed-point sums: the same bits on every run
    F = F.cpu().numpy()
    for b in range(B):
        e = oracle.rel_l2(F[b], oracle.ndft1d_type1(d[b].astype(np.complex128), x[b], N))
        assert e < TOL, (b, e)
        assert oracle.rel_l2(F[b], oracle.nfft2_1d(d[b].astype(np.complex128), x[b], N, round_mode=0)) < TOL


def test_matlab_style_calls_and_edges():
    """NFFT1_1D / NFFT2_1D on single lines (1-D tensors) and empty sets; size refusals."""
    rng = np.random.default_rng(5)
    N, M = 128, 700
    x, f, d = lines(rng, 1, N, M)
    y = m.NFFT1_1D(T(f[0]), T(x[0])).cpu().numpy()
    assert y.shape == (M,)
    assert oracle.rel_l2(y, oracle.ndft1d_type2(f[0].astype(np.complex128), x[0])) < TOL
    F = m.NFFT2_1D(T(d[0]), T(x[0]), N).cpu().numpy()
    assert F.shape == (N,)
    assert oracle.rel_l2(F, oracle.ndft1d_type1(d[0].astype(np.complex128), x[0], N)) < TOL
    plan = m.Plan1D(N, device=DEV)
    e = torch.empty(3, 0, dtype=torch.float32, device=DEV)
    assert plan.type2(T(f[[0, 0, 0]]), e).shape == (3, 0)
    z = plan.type1(torch.empty(3, 0, dtype=torch.complex64, device=DEV), e)
    assert z.shape == (3, N) and torch.count_nonzero(z).item() == 0
    for bad in (7, 0, 4098, 3000):
        with pytest.raises(m.NFFT2DError):
            m.Plan1D(bad, device=DEV)
    with pytest.raises(m.NFFT2DError, match="2048"):
        m.Plan1D(4096, device=DEV).type1(T(d), T(x))


def test_backprojection_shape():
    """The per-pulse type-2 of BP_FBP_FFBP/BP_NFFT.m:18 as one batch: 64 pulses of 424 frequency
    samples evaluated at 65,536 pixel ranges each; 4 pulses x 2048 outputs against the direct sum."""
    rng = np.random.default_rng(6)
    B, N, M = 64, 424, 65536
    x = (rng.uniform(-0.45, 0.45, (B, M)) * N).astype(np.float32)
    f = ((rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))) / np.sqrt(2)).astype(np.complex64)
    print(f"case 1-D backprojection lines={B} N={N} M={M}", flush=True)
    y = m.Plan1D(N, device=DEV).type2(T(f), T(x)).cpu().numpy()
    for b in rng.choice(B, 4, replace=False):
        sel = rng.choice(M, 2048, replace=False)
        assert oracle.rel_l2(y[b][sel], oracle.ndft1d_type2(f[b].astype(np.complex128), x[b][sel])) < TOL
