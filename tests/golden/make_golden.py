"""Generate the committed golden vectors for the 2D NFFT path.

Every expected value here comes from the mathematics, computed independently
of the oracle and of the HIP library: numpy FP64 direct sums, 50-digit mpmath
evaluations of the window closed forms, and a numpy restatement of the
reference's one 2-D known-answer relation
(NFFT/NFFT2/Matlab/Test_NFFT2_2D.m:24-29, NDFT2_2D.m:1-16).

The reference holds no golden vectors, fixtures or seeded tests for this path
(its MATLAB scripts use unseeded rand and print the error only), so these
seeded vectors stand in for them.  Coordinates are drawn in FP64 and rounded to
fp32 before the expected values are computed, so every consumer sees exactly
the same inputs.

Run:  python tests/golden/make_golden.py      (writes *.npz next to this file)
"""
from __future__ import annotations

import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 20261202
K = 6
C = 2
mp.mp.dps = 50
ALPHA_MP = (2 - mp.mpf(1) / C) * mp.pi - mp.mpf("0.01")


def cn01(rng, shape):
    """Complex normal CN(0,1): Re, Im i.i.d. N(0, 1/2)."""
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0))


def cis_neg(t):
    """exp(-2 pi i t), phase reduced mod 1 first."""
    t = t - np.floor(t)
    return np.exp(-2j * np.pi * t)


def direct_type2(f, x1, x2):
    """y_l = sum_k f[k1][k2] exp(-2 pi i (k1 x1_l/N1 + k2 x2_l/N2)), k centred; also dy/dx."""
    N1, N2 = f.shape
    k1 = np.arange(-N1 // 2, N1 // 2, dtype=np.float64)
    k2 = np.arange(-N2 // 2, N2 // 2, dtype=np.float64)
    x1 = np.asarray(x1, np.float64)
    x2 = np.asarray(x2, np.float64)
    y = np.empty(len(x1), np.complex128)
    d1 = np.empty_like(y)
    d2 = np.empty_like(y)
    for s in range(0, len(x1), 2048):
        A = cis_neg(np.outer(x1[s:s + 2048], k1) / N1)
        B = cis_neg(np.outer(x2[s:s + 2048], k2) / N2)
        G = B @ f.T                                   # (m, N1): sum_k2 f[k1,k2] B[m,k2]
        Gk = (B * k2) @ f.T
        y[s:s + 2048] = np.sum(A * G, axis=1)
        d1[s:s + 2048] = np.sum(A * G * k1, axis=1) * (-2j * np.pi / N1)
        d2[s:s + 2048] = np.sum(A * Gk, axis=1) * (-2j * np.pi / N2)
    return y, d1, d2


def direct_type1(d, x1, x2, N1, N2):
    """F[k1][k2] = sum_l d_l exp(-2 pi i (k1 x1_l/N1 + k2 x2_l/N2))."""
    k1 = np.arange(-N1 // 2, N1 // 2, dtype=np.float64)
    k2 = np.arange(-N2 // 2, N2 // 2, dtype=np.float64)
    F = np.zeros((N1, N2), np.complex128)
    for s in range(0, len(x1), 2048):
        A = cis_neg(np.outer(np.asarray(x1[s:s + 2048], np.float64), k1) / N1)
        B = cis_neg(np.outer(np.asarray(x2[s:s + 2048], np.float64), k2) / N2)
        F += A.T @ (d[s:s + 2048, None] * B)
    return F


def window_mp(d):
    """phi_hat(d) = (1/pi) sinh(alpha sqrt(K^2-d^2)) / sqrt(K^2-d^2) at 50 digits."""
    d = mp.mpf(d)
    P = K * K - d * d
    if P == 0:
        return ALPHA_MP / mp.pi
    if P > 0:
        s = mp.sqrt(P)
        return mp.sinh(ALPHA_MP * s) / (s * mp.pi)
    s = mp.sqrt(-P)
    return mp.sin(ALPHA_MP * s) / (s * mp.pi)


def special_points(N1, N2, rng, M):
    """Uniform points plus the edge cases: ties of c*x, wrap at +-N/2, grid nodes,
    points outside one period."""
    x1 = (rng.random(M) - 0.5) * N1
    x2 = (rng.random(M) - 0.5) * N2
    extra1 = [0.25, -0.25, 0.75, 1.25, -N1 / 2, N1 / 2 - 0.25, 0.0, 1.0, -3.0, N1 / 2 + 1.3,
              -N1 / 2 - 2.6, 1.5 * N1, 0.5, -0.5]
    extra2 = [0.25, 0.75, -0.25, -1.25, -N2 / 2, -N2 / 2 + 0.25, 0.0, -2.0, 2.5, -N2 / 2 - 0.7,
              N2 / 2 + 3.1, -1.25 * N2, 0.5, -0.5]
    x1 = np.concatenate([x1, extra1])
    x2 = np.concatenate([x2, extra2])
    return x1.astype(np.float32), x2.astype(np.float32)


def make_tiny(rng):
    cases = {}
    for tag, (N1, N2, M) in {"a": (8, 8, 37), "b": (16, 12, 300), "c": (32, 64, 1000)}.items():
        f = cn01(rng, (N1, N2)).astype(np.complex64)
        x1, x2 = special_points(N1, N2, rng, M)
        d = cn01(rng, len(x1)).astype(np.complex64)
        y, dy1, dy2 = direct_type2(f.astype(np.complex128), x1, x2)
        F = direct_type1(d.astype(np.complex128), x1, x2, N1, N2)
        cases.update({f"{tag}_f": f, f"{tag}_x1": x1, f"{tag}_x2": x2, f"{tag}_d": d,
                      f"{tag}_y": y, f"{tag}_dy1": dy1, f"{tag}_dy2": dy2, f"{tag}_F": F})
    np.savez_compressed(os.path.join(HERE, "tiny.npz"), **cases)


def make_c1(rng):
    """Plain type-2 at N1=N2=128, M=16384 uniform in [-1/2,1/2)^2 (Potts units,
    rounded to fp32, then x_grid = N * x_potts, exact for N a power of two)."""
    N = 128
    M = 16384
    f = cn01(rng, (N, N)).astype(np.complex64)
    p1 = (rng.random(M) - 0.5).astype(np.float32)
    p2 = (rng.random(M) - 0.5).astype(np.float32)
    x1 = (p1 * np.float32(N)).astype(np.float32)
    x2 = (p2 * np.float32(N)).astype(np.float32)
    d = cn01(rng, M).astype(np.complex64)
    y, _, _ = direct_type2(f.astype(np.complex128), x1, x2)
    F = direct_type1(d.astype(np.complex128), x1, x2, N, N)
    np.savez_compressed(os.path.join(HERE, "c1_128_16384.npz"), f=f, x1=x1, x2=x2, d=d, y=y, F=F)


def make_kat(rng):
    """NFFT/NFFT2/Matlab/Test_NFFT2_2D.m:6-29, seeded: 11x11 = 121 points in
    [-2.75, 2.75)^2, N = M = 22 outputs.  transf_1 = NDFT2_2D(data, x, y, N, M)
    restated from NDFT2_2D.m (meshgrid / column-major reshape semantics)."""
    Mx = My = 11
    N = Mm = 22
    lam = 1.0
    x = (Mx * (lam / 2) * (rng.random(Mx * My) - 0.5)).astype(np.float32)
    y = (My * (lam / 2) * (rng.random(Mx * My) - 0.5)).astype(np.float32)
    data = ((rng.random(Mx * My) - 0.5) + 1j * (rng.random(Mx * My) - 0.5)).astype(np.complex64)
    u = np.arange(-N // 2, N // 2, dtype=np.float64)
    v = np.arange(-Mm // 2, Mm // 2, dtype=np.float64)
    # [U,V] = meshgrid(u,v): U(i,j) = u(j), V(i,j) = v(i), size Mm x N; reshape(.,N*Mm,1) is column-major
    U = np.tile(u[None, :], (Mm, 1)).reshape(-1, order="F")
    V = np.tile(v[:, None], (1, N)).reshape(-1, order="F")
    kern = np.exp(-1j * 2 * np.pi * np.outer(U, x.astype(np.float64)) / N) * \
        np.exp(-1j * 2 * np.pi * np.outer(V, y.astype(np.float64)) / N)
    transf1 = (kern @ data.astype(np.complex128)).reshape((N, Mm), order="F")
    np.savez_compressed(os.path.join(HERE, "kat_test_nfft2_2d.npz"), x=x, y=y, data=data,
                        transf1=transf1, N=N, M=Mm)


def make_ndft(rng):
    """NFFT/NDFT2_2D.cu with arbitrary (non-grid) u, v: out[n] = sum_m in[m]
    exp(-2 pi i (u_n X_m / Nu + v_n Y_m / Nv)), plus a grid (u, v) case as its
    only caller builds it (Synthesis/PSO/Synthesis.cu:80-95)."""
    out = {}
    M, N, Nu, Nv = 57, 41, 20, 16
    X = rng.uniform(-10, 10, M)
    Y = rng.uniform(-8, 8, M)
    u = rng.uniform(-10, 10, N)
    v = rng.uniform(-8, 8, N)
    din = cn01(rng, M)
    ker = cis_neg(np.outer(u, X) / Nu + np.outer(v, Y) / Nv)
    out.update(dict(g_X=X, g_Y=Y, g_u=u, g_v=v, g_in=din, g_out=ker @ din, g_Nu=Nu, g_Nv=Nv))
    # grid case: Nu = Nv = 8 (u in [-Nu, Nu), 2Nu x 2Nv grid, meshgrid order of Synthesis.cu:88)
    Nu = Nv = 8
    M = 100
    X = rng.uniform(-Nu, Nu, M)
    Y = rng.uniform(-Nv, Nv, M)
    uu = np.arange(-Nu, Nu, dtype=np.float64)
    vv = np.arange(-Nv, Nv, dtype=np.float64)
    U, V = np.meshgrid(uu, vv)        # (2Nv x 2Nu), U varies along columns
    u = U.reshape(-1)
    v = V.reshape(-1)
    din = cn01(rng, M)
    ker = cis_neg(np.outer(u, X) / (2 * Nu) + np.outer(v, Y) / (2 * Nv))
    out.update(dict(m_X=X, m_Y=Y, m_u=u, m_v=v, m_in=din, m_out=ker @ din, m_Nu=2 * Nu, m_Nv=2 * Nv))
    np.savez_compressed(os.path.join(HERE, "ndft.npz"), **out)


def make_window():
    """Window, derivative and I0 at 50 digits (rounded to float64)."""
    d = np.concatenate([np.linspace(-6.5, 6.5, 53), [0.0, 5.999999, 6.0, 6.000001, -6.0, 0.123456789,
                                                     3.3, 5.5, 6.25]])
    phi = np.array([float(window_mp(v)) for v in d])
    dphi = np.array([float(mp.diff(lambda t: window_mp(t), mp.mpf(v))) for v in d])
    xs = np.array([0.0, 1e-3, 0.5, 1.0, 5.0, 10.0, 20.0, 28.214333882308138, 30.0])
    i0 = np.array([float(mp.besseli(0, mp.mpf(v))) for v in xs])
    # spectral window at N = 16, 22, 128 (phi(xi_k) = I0(K sqrt(alpha^2 - xi_k^2)), xi_k = 2 pi k/(cN))
    spec = {}
    for N in (16, 22, 128):
        ks = range(-N // 2, N // 2)
        spec[f"spec_{N}"] = np.array([float(mp.besseli(0, K * mp.sqrt(ALPHA_MP ** 2 - (2 * mp.pi * k / (C * N)) ** 2)))
                                      for k in ks])
    np.savez_compressed(os.path.join(HERE, "window_mp.npz"), d=d, phi=phi, dphi=dphi, i0_x=xs,
                        i0=i0, alpha=float(ALPHA_MP), **spec)


def main():
    ss = np.random.SeedSequence(SEED)
    r_tiny, r_c1, r_kat, r_ndft = [np.random.default_rng(s) for s in ss.spawn(4)]
    make_window()
    make_tiny(r_tiny)
    make_c1(r_c1)
    make_kat(r_kat)
    make_ndft(r_ndft)
    for fn in sorted(os.listdir(HERE)):
        if fn.endswith(".npz"):
            print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    main()
