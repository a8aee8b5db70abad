"""Every device allocation of the library, made to fail in turn (nfft2d_debug_fail_alloc).

For each scenario (the plan API, batches, the NDFT plan, 1-D plans, the
drop-ins, the type-3 NUFFT) the test counts the allocations of a clean run,
then for each of them reruns the scenario with that one allocation failing.
Each such run must stop with NFFT2D_ERR_ALLOC (-4) and a message, leave no
half-built object behind, and the next clean run must give the same results
as the first.  At the end the device memory in use must be what it was after
the first clean runs (no leak on any error path).
"""
import gc

import numpy as np
import pytest
import torch

import miscellaneous_amd as m
from miscellaneous_amd import _lib

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
ERR_ALLOC = -4


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def cn(rng, n, dt=np.complex64):
    return ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)).astype(dt)


def plan_api():
    rng = np.random.default_rng(1)
    N, M = 128, 20000
    x = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
    y = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
    f = T(cn(rng, N * N).reshape(N, N))
    d = T(cn(rng, M))
    p = m.Plan(N, N, torch.complex64, device=DEV).setpts(x, y)
    a = p.type2(f)
    b = p.type1(d)
    r, gx1, gx2 = p.type2_lsq(f, d)
    c = p.lsq_adjoint()
    torch.cuda.synchronize()
    return [a, b, r, c]


def batch_api():
    rng = np.random.default_rng(2)
    sizes = [(64, 64, 3000), (128, 96, 5000), (32, 32, 500)]
    xs = [T((rng.uniform(-0.5, 0.5, s[2]) * s[0]).astype(np.float32)) for s in sizes]
    ys = [T((rng.uniform(-0.5, 0.5, s[2]) * s[1]).astype(np.float32)) for s in sizes]
    fs = [T(cn(rng, s[0] * s[1])) for s in sizes]
    ds = [T(cn(rng, s[2])) for s in sizes]
    out2 = [torch.empty(s[2], dtype=torch.complex64, device=DEV) for s in sizes]
    out1 = [torch.empty(s[0] * s[1], dtype=torch.complex64, device=DEV) for s in sizes]
    b = m.Batch(sizes, device=DEV).set_points(xs, ys).set_io(fs, out2)
    b.setpts_type2()
    b.set_io_type1(ds, out1)
    b.type1()
    torch.cuda.synchronize()
    return out2 + out1


def ndft_api():
    rng = np.random.default_rng(3)
    N, M = 64, 3000
    f = T(cn(rng, N * N).reshape(N, N))
    x = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
    y = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
    out = m.NdftPlan(N, N, "fp32", device=DEV).set_modes(f).eval(x, y)
    torch.cuda.synchronize()
    return [out]


def plan1d_api():
    rng = np.random.default_rng(4)
    N, M, B = 96, 4000, 3
    x = T((rng.uniform(-0.5, 0.5, (B, M)) * N).astype(np.float32))
    f = T(cn(rng, B * N).reshape(B, N))
    p = m.Plan1D(N, device=DEV)
    out = p.type2(f, x)
    torch.cuda.synchronize()
    return [out]


def dropins(size):
    def run():
        rng = np.random.default_rng(5)
        N, M = size, 3000
        x = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
        y = T((rng.uniform(-0.5, 0.5, M) * N).astype(np.float32))
        d = T(cn(rng, M))
        res = torch.empty(N * N, dtype=torch.complex64, device=DEV)
        m.NFFT2_2D_GPU(res, d, x, y, N, N, M)
        torch.cuda.synchronize()
        m.Destroy_cuFFT_plan_NFFT2_2D()   # the cached plan goes, so every run allocates
        return [res]
    return run


def type3_api():
    rng = np.random.default_rng(6)
    n = 2000
    x, y = rng.uniform(-3, 3, n), rng.uniform(-2, 2, n)
    s, t = rng.uniform(-5, 5, n), rng.uniform(-4, 4, n)
    f = cn(rng, n, np.complex128)
    return [torch.from_numpy(m.NUFFT_2D_host(x, y, s, t, f, 1e-9))]


SCENARIOS = {"plan": plan_api, "batch": batch_api, "ndft": ndft_api, "plan1d": plan1d_api,
             "dropin": dropins(76), "type3": type3_api}


def same(a, b):
    return all(torch.equal(u.cpu(), v.cpu()) for u, v in zip(a, b))


def close(a, b):
    """rel-L2 <= 1e-5 per output: the type-1 sums follow the arrival order of the points in a tile."""
    return all(((u.cpu() - v.cpu()).norm() <= 1e-5 * v.cpu().norm()).item() for u, v in zip(a, b))


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_every_allocation_failure(name):
    lib = _lib.load_library()
    scn = SCENARIOS[name]
    lib.nfft2d_debug_fail_alloc(-1)
    ref = scn()                                     # warm-up: caches (type-3 workspace) grow here
    gc.collect()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(DEV)[0]
    c0 = lib.nfft2d_debug_alloc_count()
    ref = scn()
    n = lib.nfft2d_debug_alloc_count() - c0
    print(f"case alloc failures {name}: {n} allocations per run", flush=True)
    failed = 0
    for i in range(n):
        lib.nfft2d_debug_fail_alloc(i)
        try:
            scn()
        except m.NFFT2DError as e:
            assert e.code == ERR_ALLOC, f"allocation {i}: {e}"
            assert "fail" in str(e) or "hipMalloc" in str(e) or "bytes" in str(e), str(e)
            failed += 1
        finally:
            lib.nfft2d_debug_fail_alloc(-1)
        gc.collect()
        again = scn()
        # the type-1 directions sum in arrival order inside a tile unit: equal to rounding
        assert same(again, ref) or close(again, ref), f"{name}: results changed after failing allocation {i}"
    assert failed == n, f"{name}: {n - failed} of {n} failed allocations went unreported"
    gc.collect()
    torch.cuda.synchronize()
    leak = free0 - torch.cuda.mem_get_info(DEV)[0]
    print(f"  {failed} failures reported; device memory change {leak / 2**20:.2f} MiB")
    assert leak <= 4 << 20
