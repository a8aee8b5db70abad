"""The reference's own native calls, compiled as C++ against include/nfft2d_overloads.hpp.

tests/cpp/reference_calls (built by __graft_entry__.build()) issues the calls of
Synthesis/PSO/Synthesis.cu:395,397,506-510 and PSO.cu:216 with their argument
order and sizes (Nu = Nv = 44: an 88 x 88 far field from M = 44 x 44 = 1936
elements): NFFT2_2D_GPU on float2 and on double2 (the overloads of
NFFT/NFFT2_2D.cuh:8-11) and NDFT2_2D_GPU with a handle argument on the
integer meshgrid (u, v = -Nu .. Nu-1).  Its inputs and outputs are checked
here against FP64 direct sums of the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "reference_calls")


def test_reference_call_syntax(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/reference_calls is not built: run __graft_entry__.build()"
    out = tmp_path / "calls.bin"
    print("case reference_calls 88x88 M=1936", flush=True)
    res = subprocess.run([BIN, str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    raw = out.read_bytes()
    Nu, Nv, M, N = np.frombuffer(raw[:16], dtype=np.int32)
    off = 16

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    X, Y = take(np.float64, M), take(np.float64, M)
    exc = take(np.complex128, M)
    far_f, far_d, far_n = take(np.complex64, N), take(np.complex128, N), take(np.complex128, N)
    assert off == len(raw)
    N1, N2 = 2 * Nv, 2 * Nu
    # NFFT2_2D_GPU(result, data, X, Y, 2 Nv, 2 Nu, M): result[k1][k2] = sum_m data_m e^{-2 pi i (k1 X_m / N1 + k2 Y_m / N2)}
    ref = oracle.ndft_type1(exc, X, Y, N1, N2).ravel()
    e_f = oracle.rel_l2(far_f, ref)
    e_d = oracle.rel_l2(far_d, ref)
    # NDFT2_2D_GPU on the meshgrid U[n] = -Nu + n / (2 Nv), V[n] = -Nv + n % (2 Nv)
    n = np.arange(N)
    U = (-Nu + n // (2 * Nv)).astype(np.float64)
    V = (-Nv + n % (2 * Nv)).astype(np.float64)
    ref_n = oracle.ndft2_2d(X, Y, U, V, exc, 2 * Nu, 2 * Nv)
    e_n = oracle.rel_l2(far_n, ref_n)
    print(f"NFFT2_2D_GPU float2 {e_f:.2e}, double2 {e_d:.2e}; NDFT2_2D_GPU {e_n:.2e}")
    assert e_f < 1e-5 and e_d < 1e-9 and e_n < 1e-12
    # the two directions agree on the meshgrid (the reference's #ifdef NFFT / NDFT switch, Synthesis.cu:394-398)
    assert oracle.rel_l2(far_d, far_n) < 1e-9


def test_type3_mex_call(tmp_path):
    """NFFT/NFFT3/C/NFFT3_2D_C_mex.cpp:65 with NUFFT_2D_CPU replaced by NUFFT_2D_GPU_host: the same host
    arrays and fftw_complex buffers; F against the oracle's restatement of NUFFT_2D_CPU and the direct sum."""
    assert os.path.exists(BIN), "tests/cpp/reference_calls is not built: run __graft_entry__.build()"
    out = tmp_path / "type3.bin"
    print("case reference_calls --type3 3000 -> 2000", flush=True)
    res = subprocess.run([BIN, "--type3", str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    raw = out.read_bytes()
    n_in, n_out = np.frombuffer(raw[:8], dtype=np.int32)
    off = 8

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    x, y, s, t = take(np.float64, n_in), take(np.float64, n_in), take(np.float64, n_out), take(np.float64, n_out)
    f, F = take(np.complex128, n_in), take(np.complex128, n_out)
    eps = take(np.float64, 1)[0]
    assert off == len(raw)
    R, _ = oracle.nufft3_2d(x, y, s, t, f, eps)
    e_r = oracle.rel_l2(F, R)
    e_d = oracle.rel_l2(F, oracle.ndft3_2d(x, y, s, t, f))
    print(f"NUFFT_2D_GPU_host vs restatement {e_r:.2e}, vs direct sum {e_d:.2e}")
    assert e_r < 1e-11 and e_d < 10 * eps
