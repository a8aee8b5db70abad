"""GPU parity of the HIP 2D NFFT against the oracle and the golden vectors.

Tolerances (rel-L2): FP32 paths 1e-5 against FP64 direct sums (the target of
the reference path; measured ~3e-7), FP64 paths 1e-9.  Coordinates in every
fixture are fp32 values, so the FP64 expected values see exactly the inputs
the FP32 kernels see.
"""
import ctypes

import numpy as np
import pytest
import torch

import miscellaneous_amd as m
import oracle

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-9
DEV = torch.device("cuda:0")


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def cn01(rng, shape, dt=np.complex64):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)).astype(dt)


def polar_points(Npix=512, n_ang=512, n_rad=2048):
    """The polar grid of the backprojection config: angles a*pi/n_ang, radii -1/2 + j/n_rad."""
    th = np.arange(n_ang) * np.pi / n_ang
    r = -0.5 + np.arange(n_rad) / n_rad
    R, TH = np.meshgrid(r, th)
    x1 = (R * np.cos(TH)).astype(np.float32).ravel() * np.float32(Npix)
    x2 = (R * np.sin(TH)).astype(np.float32).ravel() * np.float32(Npix)
    return x1.astype(np.float32), x2.astype(np.float32)


def run_type2(f, x1, x2, dtype=torch.complex64):
    N1, N2 = f.shape
    rt = torch.float32 if dtype == torch.complex64 else torch.float64
    plan = m.Plan(N1, N2, dtype, device=DEV)
    plan.setpts(T(x1).to(rt), T(x2).to(rt))
    y = plan.type2(T(f).to(dtype))
    return y.cpu().numpy(), plan


def run_type1(d, x1, x2, N1, N2, dtype=torch.complex64):
    rt = torch.float32 if dtype == torch.complex64 else torch.float64
    plan = m.Plan(N1, N2, dtype, device=DEV)
    plan.setpts(T(x1).to(rt), T(x2).to(rt))
    return plan.type1(T(d).to(dtype)).cpu().numpy()


# ------------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_tiny_type2_type1_fp32(golden, tag):
    g = golden("tiny.npz")
    f = g[f"{tag}_f"]
    N1, N2 = f.shape
    print(f"case tiny-{tag} N={N1}x{N2} M={len(g[f'{tag}_x1'])}", flush=True)
    y, _ = run_type2(f, g[f"{tag}_x1"], g[f"{tag}_x2"])
    assert oracle.rel_l2(y, g[f"{tag}_y"]) < TOL32
    F = run_type1(g[f"{tag}_d"], g[f"{tag}_x1"], g[f"{tag}_x2"], N1, N2)
    assert oracle.rel_l2(F, g[f"{tag}_F"]) < TOL32


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_tiny_type2_type1_fp64(golden, tag):
    g = golden("tiny.npz")
    f = g[f"{tag}_f"]
    N1, N2 = f.shape
    y, _ = run_type2(f.astype(np.complex128), g[f"{tag}_x1"].astype(np.float64), g[f"{tag}_x2"].astype(np.float64),
                     torch.complex128)
    assert oracle.rel_l2(y, g[f"{tag}_y"]) < TOL64
    F = run_type1(g[f"{tag}_d"].astype(np.complex128), g[f"{tag}_x1"].astype(np.float64),
                  g[f"{tag}_x2"].astype(np.float64), N1, N2, torch.complex128)
    assert oracle.rel_l2(F, g[f"{tag}_F"]) < TOL64


def test_c1_128_type2_type1(golden):
    g = golden("c1_128_16384.npz")
    print("case c1 128x128 M=16384", flush=True)
    y, _ = run_type2(g["f"], g["x1"], g["x2"])
    e2 = oracle.rel_l2(y, g["y"])
    F = run_type1(g["d"], g["x1"], g["x2"], 128, 128)
    e1 = oracle.rel_l2(F, g["F"])
    print(f"c1 rel-L2 type-2 {e2:.3e} type-1 {e1:.3e}")
    assert e2 < TOL32 and e1 < TOL32
    # and against the oracle's FP32 restatement of the reference algorithm
    assert oracle.rel_l2(y, oracle.nfft1_2d(g["f"], g["x1"], g["x2"], precision="float")) < TOL32


def test_reference_kat_relation_dropin(golden):
    """Test_NFFT2_2D.m:24-29 through the drop-in NFFT2_2D_GPU (N = 22: non-power-of-two path)."""
    g = golden("kat_test_nfft2_2d.npz")
    N, Mm = int(g["N"]), int(g["M"])
    for dt, rt in ((torch.complex64, torch.float32), (torch.complex128, torch.float64)):
        data = T(g["data"]).to(dt)
        out = torch.empty(Mm * N, dtype=dt, device=DEV)
        m.NFFT2_2D_GPU(out, data, T(g["y"]).to(rt), T(g["x"]).to(rt), Mm, N, len(g["x"]))
        torch.cuda.synchronize()
        err = 100 * oracle.rel_l2(out.cpu().numpy().reshape(Mm, N), g["transf1"])
        assert err < (1e-3 if dt == torch.complex64 else 1e-7), err


def test_ndft2_2d_dropin(golden):
    g = golden("ndft.npz")
    for p in ("g", "m"):
        out = torch.empty(len(g[f"{p}_u"]), dtype=torch.complex128, device=DEV)
        m.NDFT2_2D_GPU(0, T(g[f"{p}_X"]), T(g[f"{p}_Y"]), T(g[f"{p}_u"]), T(g[f"{p}_v"]), T(g[f"{p}_in"]), out,
                       int(g[f"{p}_Nu"]), int(g[f"{p}_Nv"]), len(g[f"{p}_X"]), len(g[f"{p}_u"]))
        torch.cuda.synchronize()
        assert oracle.rel_l2(out.cpu().numpy(), g[f"{p}_out"]) < 1e-12


# ------------------------------------------------------------------ full-size configs
def test_polar_512_type2_and_adjoint_subsets():
    """512^2 modes, M = 1,048,576 polar points: type-2 on 4096 seeded points and
    type-1 on 1024 seeded modes against FP64 direct sums."""
    rng = np.random.default_rng(2)
    N = 512
    x1, x2 = polar_points(N)
    M = len(x1)
    print(f"case polar N={N} M={M}", flush=True)
    f = cn01(rng, (N, N))
    y, plan = run_type2(f, x1, x2)
    sel = rng.choice(M, 4096, replace=False)
    y_ref = oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel])
    e2 = oracle.rel_l2(y[sel], y_ref)
    d = cn01(rng, M)
    F = plan.type1(T(d)).cpu().numpy()
    i1 = rng.integers(0, N, 1024)
    i2 = rng.integers(0, N, 1024)
    F_ref = oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N, N, i1, i2)
    e1 = oracle.rel_l2(F[i1, i2], F_ref)
    print(f"polar rel-L2 type-2 {e2:.3e} type-1 {e1:.3e}")
    assert e2 < TOL32 and e1 < TOL32


def test_1024_large_type2_properties():
    """1024^2 modes, M = 4,194,304 uniform points: subset parity, linearity and
    the adjoint identity at full size."""
    rng = np.random.default_rng(3)
    N, M = 1024, 4 * 1024 * 1024
    print(f"case uniform N={N} M={M}", flush=True)
    x1 = ((rng.random(M) - 0.5).astype(np.float32) * np.float32(N)).astype(np.float32)
    x2 = ((rng.random(M) - 0.5).astype(np.float32) * np.float32(N)).astype(np.float32)
    f = cn01(rng, (N, N))
    g = cn01(rng, (N, N))
    plan = m.Plan(N, N, device=DEV).setpts(T(x1), T(x2))
    ft, gt = T(f), T(g)
    yf = plan.type2(ft)
    yg = plan.type2(gt)
    ylin = plan.type2(ft * (0.5 - 2j) + gt * 1.5)
    lin = (ylin - (yf * (0.5 - 2j) + yg * 1.5)).norm() / ylin.norm()
    assert lin.item() < 1e-6
    sel = rng.choice(M, 2048, replace=False)
    y_ref = oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel])
    assert oracle.rel_l2(yf.cpu().numpy()[sel], y_ref) < TOL32
    r = T(cn01(rng, M))
    lhs = torch.vdot(r.to(torch.complex128), yf.to(torch.complex128))
    rhs = torch.vdot(plan.adjoint(r).to(torch.complex128).ravel(), ft.to(torch.complex128).ravel())
    assert (abs(lhs - rhs) / abs(lhs)).item() < 1e-5


# ------------------------------------------------------------------ edge cases
def test_empty_and_single_point():
    N = 32
    plan = m.Plan(N, N, device=DEV)
    plan.setpts(torch.empty(0, device=DEV), torch.empty(0, device=DEV))
    F = plan.type1(torch.empty(0, dtype=torch.complex64, device=DEV))
    assert torch.count_nonzero(F).item() == 0
    y = plan.type2(torch.ones(N, N, dtype=torch.complex64, device=DEV))
    assert y.numel() == 0
    x, yy = 3.37, -4.81
    plan.setpts(torch.tensor([x], device=DEV), torch.tensor([yy], device=DEV))
    F = plan.type1(torch.ones(1, dtype=torch.complex64, device=DEV)).cpu().numpy()
    k = np.arange(-N // 2, N // 2)
    ref = np.exp(-2j * np.pi * (k[:, None] * np.float32(x) / N + k[None, :] * np.float32(yy) / N))
    assert oracle.rel_l2(F, ref) < TOL32


def test_plan_state_and_size_errors():
    """Host-side refusals of a live plan: no points yet, a point count above
    2^30, set_maxpts resetting the points.  No kernel is launched by any of the
    refused calls."""
    plan = m.Plan(32, 32, device=DEV)
    lib, h = plan._lib, plan._h
    assert lib.nfft2d_plan_num_points(h) == -5
    assert "no points" in lib.nfft2d_last_error().decode()
    out = torch.empty(4, dtype=torch.complex64, device=DEV)
    f = torch.zeros(32, 32, dtype=torch.complex64, device=DEV)
    assert lib.nfft2d_type2(h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(f.data_ptr()), None) == -5
    x = torch.zeros(4, device=DEV)
    assert lib.nfft2d_setpts(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()), (1 << 30) + 1,
                             None) == -1
    assert "outside" in lib.nfft2d_last_error().decode()
    plan.setpts(x, x)
    assert lib.nfft2d_plan_num_points(h) == 4
    assert lib.nfft2d_plan_set_maxpts(h, 64) == 0
    assert lib.nfft2d_plan_num_points(h) == -5


def test_lazy_conjugate_views_are_refused_or_resolved():
    """torch's .conj() only sets a flag; the library reads memory, so the plan API
    refuses such views, and the autograd wrapper resolves them."""
    rng = np.random.default_rng(17)
    N, M = 32, 500
    x1 = T(((rng.random(M) - 0.5) * N).astype(np.float32))
    x2 = T(((rng.random(M) - 0.5) * N).astype(np.float32))
    f = T(cn01(rng, (N, N)))
    plan = m.Plan(N, N, device=DEV).setpts(x1, x2)
    with pytest.raises(ValueError, match="lazy conjugate"):
        plan.type2(f.conj())
    y = m.nfft2d_type2(f.conj(), x1, x2, plan)
    assert torch.equal(y, plan.type2(f.conj().resolve_conj()))


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_clustered_points_and_dense_tiles(dtype):
    """All points in a few cells (many subproblems per tile, so the spread's
    two-subproblem units and their second pass) and exact ties."""
    rng = np.random.default_rng(4)
    N, M = 64, 50000
    x1 = np.concatenate([rng.normal(0, 0.3, M // 2), np.full(M // 4, 0.25), rng.uniform(-32, 32, M - M // 2 - M // 4)])
    x2 = np.concatenate([rng.normal(5, 0.2, M // 2), np.full(M // 4, -0.75), rng.uniform(-32, 32, M - M // 2 - M // 4)])
    x1 = x1.astype(np.float32)
    x2 = x2.astype(np.float32)
    f = cn01(rng, (N, N))
    d = cn01(rng, M)
    wide = dtype == torch.complex128
    tol = TOL64 if wide else TOL32
    y, plan = run_type2(f.astype(np.complex128) if wide else f, x1.astype(np.float64) if wide else x1,
                        x2.astype(np.float64) if wide else x2, dtype)
    sel = rng.choice(M, 3000, replace=False)
    assert oracle.rel_l2(y[sel], oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel])) < tol
    F = plan.type1(T(d).to(dtype)).cpu().numpy()
    assert oracle.rel_l2(F, oracle.ndft_type1(d.astype(np.complex128), x1, x2, N, N)) < tol


def test_small_last_subproblems_join_the_one_before():
    """Bins of 256 q + k points, k on both sides of 128: the complex64 gather lets
    a bin's last subproblem of at most 128 points share the block (tile and
    staging) of the full one before it.  Its values must be the same bits as
    the one-point-per-thread gather behind type2_dx, and match the direct sum."""
    rng = np.random.default_rng(12)
    N = 64                                       # 128^2 grid, 8 x 8 bins of 16 x 16 cells
    extra = [0, 1, 2, 64, 127, 128, 129, 200, 255, 256]
    xs, ys = [], []
    for b in range(64):
        n = 256 * (1 + b % 3) + extra[b % len(extra)]
        mu1 = 16 * (b % 8) + rng.integers(0, 16, n)
        mu2 = 16 * (b // 8) + rng.integers(0, 16, n)
        # cell mu (mod 128) holds the x with rint(2 x) = mu - 128 [mu >= 64]
        xs.append((np.where(mu1 >= 64, mu1 - 128, mu1) + rng.uniform(-0.45, 0.45, n)) / 2)
        ys.append((np.where(mu2 >= 64, mu2 - 128, mu2) + rng.uniform(-0.45, 0.45, n)) / 2)
    x1 = np.concatenate(xs).astype(np.float32)
    x2 = np.concatenate(ys).astype(np.float32)
    perm = rng.permutation(len(x1))
    x1, x2 = x1[perm], x2[perm]
    f = cn01(rng, (N, N))
    plan = m.Plan(N, N, device=DEV).setpts(T(x1), T(x2))
    y = plan.type2(T(f))
    y1, _, _ = plan.type2_dx(T(f))
    assert torch.equal(y, y1)
    sel = rng.choice(len(x1), 2000, replace=False)
    assert oracle.rel_l2(y.cpu().numpy()[sel], oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel])) < TOL32


@pytest.mark.parametrize("N1,N2", [(8, 8), (16, 64), (256, 32), (2048, 16), (4096, 8), (6, 10), (88, 88)])
def test_shapes(N1, N2):
    rng = np.random.default_rng(N1 * 7 + N2)
    M = 2000
    x1 = ((rng.random(M) - 0.5) * N1).astype(np.float32)
    x2 = ((rng.random(M) - 0.5) * N2).astype(np.float32)
    f = cn01(rng, (N1, N2))
    d = cn01(rng, M)
    print(f"case shape {N1}x{N2}", flush=True)
    y, plan = run_type2(f, x1, x2)
    assert oracle.rel_l2(y, oracle.ndft_type2(f.astype(np.complex128), x1, x2)) < TOL32
    F = plan.type1(T(d)).cpu().numpy()
    i1 = rng.integers(0, N1, 512)
    i2 = rng.integers(0, N2, 512)
    assert oracle.rel_l2(F[i1, i2], oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i1, i2)) < TOL32


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("N1,N2", [(26, 34), (34, 26), (42, 64), (50, 50), (88, 90)])
def test_spread_thin_last_bin(N1, N2, dtype):
    """Grids whose last 16-cell bin is 4 cells wide (2 N = 4 mod 16, i.e. N = 2 mod 8): the
    spread tiles of the bins beside it reach two bins across the wrap, so the tile merge visits
    the bins at offsets -2..2 in that dimension.  Type-1 and adjoint against FP64 direct sums,
    with points crowded near the wrap and spread over the whole period."""
    rng = np.random.default_rng(N1 * 100 + N2)
    M = 3000
    x1 = np.concatenate([rng.uniform(-N1 / 2, N1 / 2, M // 2), N1 / 2 - rng.uniform(0, 3, M - M // 2)])
    x2 = np.concatenate([rng.uniform(-N2 / 2, N2 / 2, M // 2), -N2 / 2 + rng.uniform(0, 3, M - M // 2)])
    x1 = x1.astype(np.float32)
    x2 = x2.astype(np.float32)
    wide = dtype == torch.complex128
    rt = torch.float64 if wide else torch.float32
    tol = TOL64 if wide else TOL32
    d = cn01(rng, M, np.complex128 if wide else np.complex64)
    print(f"case thin last bin {N1}x{N2} {dtype}", flush=True)
    plan = m.Plan(N1, N2, dtype, device=DEV).setpts(T(x1).to(rt), T(x2).to(rt))
    F = plan.type1(T(d)).cpu().numpy()
    Fa = plan.adjoint(T(d)).cpu().numpy()
    ref = oracle.ndft_type1(d.astype(np.complex128), x1, x2, N1, N2)
    ref_a = np.conj(oracle.ndft_type1(np.conj(d.astype(np.complex128)), x1, x2, N1, N2))
    e1, ea = oracle.rel_l2(F, ref), oracle.rel_l2(Fa, ref_a)
    print(f"thin last bin {N1}x{N2}: type-1 {e1:.2e}, adjoint {ea:.2e}")
    assert e1 < tol and ea < tol


@pytest.mark.parametrize("N1,N2", [(126, 126), (124, 22), (22, 124), (120, 96)])
def test_type1_matrix_core_line_lengths(N1, N2):
    """complex64 type-1 passes of even lengths up to 126 that are not powers of two run as
    complex matrix products on the FP32 matrix cores (16 x 16 output tiles, the K sum split
    over 4 waves): the largest such length, lopsided shapes with one short side, and 120 x 96.
    Type-1 and adjoint, every mode, against FP64 direct sums."""
    rng = np.random.default_rng(N1 * 1000 + N2)
    M = 3000
    x1 = ((rng.random(M) - 0.5) * N1).astype(np.float32)
    x2 = ((rng.random(M) - 0.5) * N2).astype(np.float32)
    d = cn01(rng, M)
    print(f"case matrix-core type-1 {N1}x{N2}", flush=True)
    plan = m.Plan(N1, N2, torch.complex64, device=DEV).setpts(T(x1), T(x2))
    F = plan.type1(T(d)).cpu().numpy()
    Fa = plan.adjoint(T(d)).cpu().numpy()
    ref = oracle.ndft_type1(d.astype(np.complex128), x1, x2, N1, N2)
    ref_a = np.conj(oracle.ndft_type1(np.conj(d.astype(np.complex128)), x1, x2, N1, N2))
    e1, ea = oracle.rel_l2(F, ref), oracle.rel_l2(Fa, ref_a)
    print(f"matrix-core type-1 {N1}x{N2}: type-1 {e1:.2e}, adjoint {ea:.2e}")
    assert e1 < TOL32 and ea < TOL32


@pytest.mark.parametrize("N1,N2,M", [(2048, 1024, 300000), (2048, 2048, 200000)])
def test_large_grids_sort_paths(N1, N2, M):
    """Oversampled grids above 16,384 bins: the point histogram goes through
    global atomics (k_hist_global); at 2048 x 1024 (32,768 bins) the group
    scan stays folded into k_scatter_scan, at 2048 x 2048 (65,536 bins) it runs
    as the one-block k_scan_groups + k_scatter.  Type-2 on 256 seeded points
    and type-1 on 256 seeded modes against FP64 direct sums, both with half
    the points clustered into a few bins (many subproblems per bin)."""
    rng = np.random.default_rng(N1 + N2)
    h = M // 2
    x1 = np.concatenate([(rng.random(h) - 0.5) * N1, rng.normal(0.0, 3.0, M - h)]).astype(np.float32)
    x2 = np.concatenate([(rng.random(h) - 0.5) * N2, rng.normal(0.0, 3.0, M - h)]).astype(np.float32)
    f = cn01(rng, (N1, N2))
    d = cn01(rng, M)
    print(f"case large grid {N1}x{N2} M={M}", flush=True)
    y, plan = run_type2(f, x1, x2)
    sel = rng.choice(M, 256, replace=False)
    e2 = oracle.rel_l2(y[sel], oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel]))
    F = plan.type1(T(d)).cpu().numpy()
    i1 = rng.integers(0, N1, 256)
    i2 = rng.integers(0, N2, 256)
    e1 = oracle.rel_l2(F[i1, i2], oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i1, i2))
    print(f"large grid rel-L2 type-2 {e2:.3e} type-1 {e1:.3e}")
    assert e2 < TOL32 and e1 < TOL32


def test_nonfinite_coordinate_is_contained():
    rng = np.random.default_rng(5)
    N, M = 32, 500
    x1 = ((rng.random(M) - 0.5) * N).astype(np.float32)
    x2 = ((rng.random(M) - 0.5) * N).astype(np.float32)
    x1[7] = np.nan
    x2[11] = np.inf
    f = cn01(rng, (N, N))
    y, _ = run_type2(f, x1, x2)
    ok = np.ones(M, bool)
    ok[[7, 11]] = False
    assert np.isnan(y[7]) and np.isnan(y[11])
    assert oracle.rel_l2(y[ok], oracle.ndft_type2(f.astype(np.complex128), x1[ok], x2[ok])) < TOL32


def test_type2_bitwise_deterministic():
    rng = np.random.default_rng(6)
    N, M = 128, 100000
    x1 = ((rng.random(M) - 0.5) * N).astype(np.float32)
    x2 = ((rng.random(M) - 0.5) * N).astype(np.float32)
    f = T(cn01(rng, (N, N)))
    plan = m.Plan(N, N, device=DEV).setpts(T(x1), T(x2))
    a = plan.type2(f)
    plan.setpts(T(x1), T(x2))
    b = plan.type2(f)
    assert torch.equal(a, b)


def test_setpts_type2_fused_matches_and_replays():
    """setpts + type-2 in one call (the tile sort on the plan's second stream,
    beside the grid passes) gives exactly what the two calls give, eagerly and
    replayed from a graph, with new coordinates written into the captured
    inputs between the replays."""
    rng = np.random.default_rng(12)
    N, M = 256, 200000

    def pts():
        return (((rng.random(M) - 0.5) * N).astype(np.float32), ((rng.random(M) - 0.5) * N).astype(np.float32))

    fn = cn01(rng, (N, N))
    f = T(fn)
    a1, a2 = pts()
    x1, x2 = T(a1), T(a2)
    plan = m.Plan(N, N, device=DEV)
    ref = m.Plan(N, N, device=DEV)
    y = plan.setpts_type2(x1, x2, f)
    torch.cuda.synchronize()
    assert torch.equal(y, ref.setpts(x1, x2).type2(f))
    sel = np.sort(rng.choice(M, 512, replace=False))
    err = oracle.rel_l2(y.cpu().numpy()[sel], oracle.ndft_type2(fn.astype(np.complex128), a1[sel], a2[sel]))
    print(f"fused setpts+type2 256^2 M={M}: rel-L2 {err:.2e}")
    assert err < TOL32
    out = torch.empty(M, dtype=torch.complex64, device=DEV)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.setpts_type2(x1, x2, f, out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.setpts_type2(x1, x2, f, out)
    for rep in range(3):
        b1, b2 = pts()
        x1.copy_(T(b1))
        x2.copy_(T(b2))
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        expect = ref.setpts(x1, x2).type2(f)
        torch.cuda.synchronize()
        assert torch.equal(out, expect), rep


def test_dropins_match_plan_api():
    rng = np.random.default_rng(8)
    N1, N2, M = 64, 128, 5000
    x1 = T(((rng.random(M) - 0.5) * N1).astype(np.float32))
    x2 = T(((rng.random(M) - 0.5) * N2).astype(np.float32))
    f = T(cn01(rng, (N1, N2)))
    d = T(cn01(rng, M))
    m.Calculate_cuFFT_plan_C2C_NFFT2_2D(N1, N2)
    out2 = torch.empty(M, dtype=torch.complex64, device=DEV)
    m.NFFT1_2D_GPU(out2, f, x1, x2, N1, N2, M)
    out1 = torch.empty(N1 * N2, dtype=torch.complex64, device=DEV)
    m.NFFT2_2D_GPU(out1, d, x1, x2, N1, N2, M)
    torch.cuda.synchronize()
    plan = m.Plan(N1, N2, device=DEV).setpts(x1, x2)
    assert torch.equal(out2, plan.type2(f))
    assert (out1.reshape(N1, N2) - plan.type1(d)).abs().max().item() < 1e-3 * out1.abs().max().item()
    m.Destroy_cuFFT_plan_NFFT2_2D()


# ------------------------------------------------------------------ gradient
def test_type2_dx_vs_analytic():
    rng = np.random.default_rng(9)
    N1, N2, M = 64, 48, 3000
    x1 = ((rng.random(M) - 0.5) * N1).astype(np.float32)
    x2 = ((rng.random(M) - 0.5) * N2).astype(np.float32)
    f = cn01(rng, (N1, N2))
    plan = m.Plan(N1, N2, device=DEV).setpts(T(x1), T(x2))
    y, d1, d2 = plan.type2_dx(T(f))
    yr, d1r, d2r = oracle.ndft_type2(f.astype(np.complex128), x1, x2, with_dx=True)
    assert oracle.rel_l2(y.cpu().numpy(), yr) < TOL32
    assert oracle.rel_l2(d1.cpu().numpy(), d1r) < 1e-5
    assert oracle.rel_l2(d2.cpu().numpy(), d2r) < 1e-5
    p64 = m.Plan(N1, N2, torch.complex128, device=DEV).setpts(T(x1).double(), T(x2).double())
    y, d1, d2 = p64.type2_dx(T(f).to(torch.complex128))
    assert oracle.rel_l2(d1.cpu().numpy(), d1r) < 1e-9
    assert oracle.rel_l2(d2.cpu().numpy(), d2r) < 1e-9


def test_grad_x_on_held_grid_matches_recompute():
    """The derivative gather on the grid the forward left in the plan equals
    the one that recomputes the grid, and the autograd backward gives the same
    gradients whether or not another transform rewrote the grid in between."""
    rng = np.random.default_rng(14)
    N1, N2, M = 64, 96, 5000
    x1 = T(((rng.random(M) - 0.5) * N1).astype(np.float32))
    x2 = T(((rng.random(M) - 0.5) * N2).astype(np.float32))
    f = T(cn01(rng, (N1, N2)))
    g = T(cn01(rng, M))
    plan = m.Plan(N1, N2, device=DEV).setpts(x1, x2)
    plan.type2(f)
    a1, a2 = plan.type2_grad_x_grid(g)
    b1, b2 = plan.type2_grad_x(g, f)
    assert torch.equal(a1, b1) and torch.equal(a2, b2)

    def grads(rewrite):
        xt = x1.clone().requires_grad_(True)
        yt = x2.clone().requires_grad_(True)
        ft = f.clone().requires_grad_(True)
        pred = m.nfft2d_type2(ft, xt, yt, plan)
        if rewrite:
            plan.type1(g)                      # overwrites the plan's grid
        (pred * g.conj()).real.sum().backward()
        return ft.grad, xt.grad, yt.grad

    held = grads(False)
    recomputed = grads(True)
    # the coordinate gradients come from per-point gathers: the same bits; the
    # mode gradient is a spread, whose summation order follows the point order
    # inside a bin (the sort is redone by every forward)
    assert torch.equal(held[1], recomputed[1]) and torch.equal(held[2], recomputed[2])
    assert ((held[0] - recomputed[0]).norm() / recomputed[0].norm()).item() < 1e-6


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_least_squares_gather_matches_the_separate_calls(dtype):
    """type2_lsq (value, residual and coordinate gradient in one gather) gives
    the bits of type2 + r = 2 (y - y_obs) + type2_grad_x_grid(r)."""
    rng = np.random.default_rng(15)
    N1, N2, M = 128, 64, 20000
    rt = torch.float32 if dtype == torch.complex64 else torch.float64
    x1 = T(((rng.random(M) - 0.5) * N1).astype(np.float32)).to(rt)
    x2 = T(((rng.random(M) - 0.5) * N2).astype(np.float32)).to(rt)
    f = T(cn01(rng, (N1, N2))).to(dtype)
    y_obs = T(cn01(rng, M)).to(dtype)
    plan = m.Plan(N1, N2, dtype, device=DEV).setpts(x1, x2)
    r, gx1, gx2 = plan.type2_lsq(f, y_obs)
    y = plan.type2(f)
    r_ref = torch.add(-2 * y_obs, y, alpha=2)
    g1, g2 = plan.type2_grad_x_grid(r_ref)
    assert torch.equal(r, r_ref)
    assert torch.equal(gx1, g1) and torch.equal(gx2, g2)
    # the residual the plan keeps: its adjoint is the bits of adjoint(r), with
    # or without r written out in point order (the MFMA spreads of both
    # precisions sum in a fixed order)
    gf = plan.adjoint(r)
    _, hx1, hx2 = plan.type2_lsq(f, y_obs, keep_r=False)
    assert torch.equal(hx1, gx1) and torch.equal(hx2, gx2)
    for _ in range(2):                                 # twice: the held residual is reusable
        assert torch.equal(plan.lsq_adjoint(), gf)


def test_lsq_adjoint_state():
    """lsq_adjoint needs a type2_lsq on the current point set: refused before
    the first one and again after setpts; an empty point set gives zeros."""
    rng = np.random.default_rng(16)
    N, M = 64, 3000
    x1 = T(((rng.random(M) - 0.5) * N).astype(np.float32))
    x2 = T(((rng.random(M) - 0.5) * N).astype(np.float32))
    plan = m.Plan(N, N, device=DEV).setpts(x1, x2)
    with pytest.raises(m.NFFT2DError, match="no residual"):
        plan.lsq_adjoint()
    f = T(cn01(rng, (N, N)))
    y_obs = T(cn01(rng, M))
    r, _, _ = plan.type2_lsq(f, y_obs)
    assert torch.equal(plan.lsq_adjoint(), plan.adjoint(r))
    plan.setpts(x2, x1)
    with pytest.raises(m.NFFT2DError, match="no residual"):
        plan.lsq_adjoint()
    plan.setpts(torch.empty(0, device=DEV), torch.empty(0, device=DEV))
    assert torch.count_nonzero(plan.lsq_adjoint()).item() == 0


def test_autograd_gradients_vs_fp64_and_finite_differences():
    """loss = ||A(x) f - y||^2 at 256^2, M = 262,144 (subset checks for dL/dx)."""
    rng = np.random.default_rng(10)
    N, M = 256, 262144
    print(f"case grad N={N} M={M}", flush=True)
    x1 = ((rng.random(M) - 0.5).astype(np.float32) * np.float32(N)).astype(np.float32)
    x2 = ((rng.random(M) - 0.5).astype(np.float32) * np.float32(N)).astype(np.float32)
    f_true = cn01(rng, (N, N))
    y_obs = oracle.nfft1_2d(f_true, x1, x2, precision="float") + \
        (np.sqrt(1e-3 / 2) * (rng.standard_normal(M) + 1j * rng.standard_normal(M))).astype(np.complex64)
    f0 = cn01(rng, (N, N))
    plan = m.Plan(N, N, device=DEV)
    ft = T(f0).requires_grad_(True)
    xt = T(x1).requires_grad_(True)
    yt = T(x2).requires_grad_(True)
    pred = m.nfft2d_type2(ft, xt, yt, plan)
    loss = ((pred - T(y_obs)).abs() ** 2).sum()
    loss.backward()
    # FP64 reference gradient w.r.t. f: 2 A^H r (residual from the FP64 restatement, ~5e-12 from the direct sum)
    r = oracle.nfft1_2d(f0.astype(np.complex128), x1, x2) - y_obs.astype(np.complex128)
    sel_modes = (rng.integers(0, N, 512), rng.integers(0, N, 512))
    gf_ref = 2 * np.conj(oracle.ndft_type1_modes(np.conj(r), x1, x2, N, N, *sel_modes))
    gf = ft.grad.cpu().numpy()[sel_modes]
    assert oracle.rel_l2(gf, gf_ref) < 1e-5
    # w.r.t. x: 2 Re(conj(r) dy/dx) on 64 seeded points, analytic FP64 and central FD (h = 1e-3 grid units)
    sel = rng.choice(M, 64, replace=False)
    _, d1, d2 = oracle.ndft_type2(f0.astype(np.complex128), x1[sel], x2[sel], with_dx=True)
    gx1_ref = 2 * np.real(np.conj(r[sel]) * d1)
    gx2_ref = 2 * np.real(np.conj(r[sel]) * d2)
    gx1 = xt.grad.cpu().numpy()[sel]
    gx2 = yt.grad.cpu().numpy()[sel]
    assert oracle.rel_l2(gx1, gx1_ref) < 1e-4
    assert oracle.rel_l2(gx2, gx2_ref) < 1e-4
    h = 1e-3
    xs1 = x1[sel].astype(np.float64)
    xs2 = x2[sel].astype(np.float64)
    fd1 = (np.abs(oracle.ndft_type2(f0.astype(np.complex128), xs1 + h, xs2) - y_obs[sel]) ** 2 -
           np.abs(oracle.ndft_type2(f0.astype(np.complex128), xs1 - h, xs2) - y_obs[sel]) ** 2) / (2 * h)
    fd2 = (np.abs(oracle.ndft_type2(f0.astype(np.complex128), xs1, xs2 + h) - y_obs[sel]) ** 2 -
           np.abs(oracle.ndft_type2(f0.astype(np.complex128), xs1, xs2 - h) - y_obs[sel]) ** 2) / (2 * h)
    e1, e2 = oracle.rel_l2(gx1, fd1), oracle.rel_l2(gx2, fd2)
    print(f"grad_x vs central differences: d/dx1 {e1:.2e}, d/dx2 {e2:.2e}")
    assert e1 < 1e-3 and e2 < 1e-3


def test_autograd_skips_the_sort_for_unchanged_points():
    """Repeated forwards at the same coordinate tensors sort once; an in-place update of the
    coordinates (an optimiser step) is seen through the tensors' version counters and re-sorts.
    Results equal a fresh plan's in both cases."""
    rng = np.random.default_rng(23)
    N, M = 64, 5000
    x1 = T(((rng.random(M) - 0.5) * N).astype(np.float32)).requires_grad_(True)
    x2 = T(((rng.random(M) - 0.5) * N).astype(np.float32)).requires_grad_(True)
    f = T(cn01(rng, (N, N))).requires_grad_(True)
    plan = m.Plan(N, N, device=DEV)
    y0 = m.nfft2d_type2(f, x1, x2, plan)
    v0 = plan.pts_version
    y1 = m.nfft2d_type2(f, x1, x2, plan)
    assert plan.pts_version == v0 and torch.equal(y0, y1)
    (y1.abs() ** 2).sum().backward()
    with torch.no_grad():
        x1.add_(0.25)                              # in place: the version counter moves
    y2 = m.nfft2d_type2(f, x1, x2, plan)
    assert plan.pts_version == v0 + 1
    fresh = m.Plan(N, N, device=DEV).setpts(x1.detach().contiguous(), x2.detach().contiguous())
    assert torch.equal(y2, fresh.type2(f.detach()))
    plan.setpts(x2.detach().contiguous(), x1.detach().contiguous())   # re-pointed elsewhere
    y3 = m.nfft2d_type2(f, x1, x2, plan)
    assert torch.equal(y3, y2)


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("N1,N2", [(1000, 1000), (130, 500), (512, 250), (2046, 128), (1998, 64)])
def test_chirp_z_sizes(N1, N2, dtype):
    """Sizes that are not powers of two, from 128 on: the FFT passes are chirp-z transforms
    (Bluestein, at the power of two >= 3N - 1).  Type-2, type-1 and the adjoint (the e^{+}
    passes) against FP64 direct sums, points over 1.5 periods."""
    rng = np.random.default_rng(N1 * 3 + N2)
    M = 40000
    x1 = (rng.uniform(-0.5, 0.5, M) * N1 * 1.5)
    x2 = (rng.uniform(-0.5, 0.5, M) * N2 * 1.5)
    cdt = np.complex64 if dtype == torch.complex64 else np.complex128
    rdt = np.float32 if dtype == torch.complex64 else np.float64
    x1, x2 = x1.astype(rdt), x2.astype(rdt)
    f = cn01(rng, (N1, N2), cdt)
    d = cn01(rng, M, cdt)
    tol = TOL32 if dtype == torch.complex64 else 1e-9
    print(f"case chirp-z {N1}x{N2} {dtype}", flush=True)
    plan = m.Plan(N1, N2, dtype, device=DEV).setpts(T(x1), T(x2))
    y = plan.type2(T(f)).cpu().numpy()
    sel = rng.choice(M, 1500, replace=False)
    e2 = oracle.rel_l2(y[sel], oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel]))
    F = plan.type1(T(d)).cpu().numpy()
    i1, i2 = rng.integers(0, N1, 400), rng.integers(0, N2, 400)
    e1 = oracle.rel_l2(F[i1, i2], oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i1, i2))
    A = plan.adjoint(T(d)).cpu().numpy()   # sum_l d_l e^{+2 pi i (k1 x1 / N1 + k2 x2 / N2)}
    ea = oracle.rel_l2(A[i1, i2], np.conj(oracle.ndft_type1_modes(np.conj(d).astype(np.complex128), x1, x2, N1,
                                                                     N2, i1, i2)))
    print(f"  rel-L2 type-2 {e2:.2e}, type-1 {e1:.2e}, adjoint {ea:.2e}")
    assert e2 < tol and e1 < tol and ea < tol


def test_chirp_z_complex64_up_to_4096():
    """Complex64 plans take every even N up to 4096 (chirp-z lines of 16,384 points at N = 3000)."""
    rng = np.random.default_rng(77)
    N1, N2, M = 3000, 130, 20000
    x1 = (rng.uniform(-0.5, 0.5, M) * N1).astype(np.float32)
    x2 = (rng.uniform(-0.5, 0.5, M) * N2).astype(np.float32)
    f = cn01(rng, (N1, N2))
    d = cn01(rng, M)
    print("case chirp-z 3000x130 complex64", flush=True)
    plan = m.Plan(N1, N2, torch.complex64, device=DEV).setpts(T(x1), T(x2))
    y = plan.type2(T(f)).cpu().numpy()
    sel = rng.choice(M, 1000, replace=False)
    assert oracle.rel_l2(y[sel], oracle.ndft_type2(f.astype(np.complex128), x1[sel], x2[sel])) < TOL32
    F = plan.type1(T(d)).cpu().numpy()
    i1, i2 = rng.integers(0, N1, 300), rng.integers(0, N2, 300)
    assert oracle.rel_l2(F[i1, i2], oracle.ndft_type1_modes(d.astype(np.complex128), x1, x2, N1, N2, i1, i2)) < TOL32

