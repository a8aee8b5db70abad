"""Type-3 2-D NUFFT (non-uniform to non-uniform, FP64) over the C ABI.

``NUFFT_2D(x, y, s, t, f, eps)`` computes
``F[j] = sum_k f[k] exp(-1j * (s[j] * x[k] + t[j] * y[k]))`` to the accuracy
eps.  It is the counterpart of ``NUFFT_2D_CPU``
(``NFFT/NFFT3/C/NFFT3_2D_C.cpp:199``; MATLAB entry ``NFFT3_2D_C_mex.cpp:18``,
called as ``[Fr, Fi] = NFFT3_2D_C(x, y, s, t, f, eps)``).  It uses the same
argument meaning and the same parameters: R = 2.2, msp from eps, and the
Mry x Mrx grid from the extents.

* ``NUFFT_2D(...)`` takes float64 / complex128 CUDA tensors and returns F on
  the same device.
* ``NUFFT_2D_host(...)`` takes numpy arrays and returns a numpy array.  It
  keeps the host-memory contract of ``NUFFT_2D_CPU``.
* ``nufft3_params(...)`` returns the derived parameters (``msp``, ``Mrx``,
  ``Mry``, ...).

There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check

__all__ = ["NUFFT_2D", "NUFFT_2D_host", "nufft3_params"]

_KEYS = ["xb", "yb", "sb", "tb", "kkx", "kky", "Dx", "Dy", "b", "t1", "fact", "msp", "Mrx", "Mry"]


def _vec(a: torch.Tensor, dtype, name: str, device=None) -> torch.Tensor:
    if not isinstance(a, torch.Tensor) or not a.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if device is not None and a.device != device:
        raise ValueError(f"{name} is on {a.device}, expected {device}")
    if a.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {a.dtype}")
    if a.is_conj() or a.is_neg():
        raise ValueError(f"{name} is a lazy conjugate or negative view: pass a.resolve_conj().resolve_neg()")
    return a.reshape(-1).contiguous()


def _ptr(a) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.data_ptr() if a.numel() else 0)


def NUFFT_2D(x: torch.Tensor, y: torch.Tensor, s: torch.Tensor, t: torch.Tensor, f: torch.Tensor,
             eps: float = 1e-6) -> torch.Tensor:
    """F[j] = sum_k f[k] e^{-i (s[j] x[k] + t[j] y[k])} (NFFT3_2D_C.cpp:199); x, y, f: len_in; s, t: len_out."""
    dev = x.device
    xs, ys = _vec(x, torch.float64, "x", dev), _vec(y, torch.float64, "y", dev)
    ss, ts = _vec(s, torch.float64, "s", dev), _vec(t, torch.float64, "t", dev)
    fs = _vec(f, torch.complex128, "f", dev)
    if ys.numel() != xs.numel() or fs.numel() != xs.numel():
        raise ValueError(f"x, y, f must have one length, got {xs.numel()}, {ys.numel()}, {fs.numel()}")
    if ts.numel() != ss.numel():
        raise ValueError(f"s and t must have one length, got {ss.numel()}, {ts.numel()}")
    F = torch.empty(ss.numel(), dtype=torch.complex128, device=dev)
    lib = _lib.load_library()
    with torch.cuda.device(dev):
        torch.cuda.current_stream(dev).synchronize()   # the library runs on the null stream
        check(lib.NUFFT_2D_GPU(_ptr(xs), _ptr(ys), _ptr(ss), _ptr(ts), _ptr(fs), _ptr(F), float(eps),
                               xs.numel(), ss.numel()), "NUFFT_2D_GPU")
        if torch.cuda.current_stream(dev).cuda_stream != 0:   # a non-blocking stream does not wait for it
            torch.cuda.synchronize(dev)
    return F


def NUFFT_2D_host(x, y, s, t, f, eps: float = 1e-6) -> np.ndarray:
    """NUFFT_2D_CPU's contract on host arrays: numpy in, numpy out (complex128)."""
    xs, ys, ss, ts = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, y, s, t))
    fs = np.ascontiguousarray(f, dtype=np.complex128).reshape(-1)
    if ys.size != xs.size or fs.size != xs.size or ts.size != ss.size:
        raise ValueError("x, y, f must have one length and s, t another")
    F = np.empty(ss.size, dtype=np.complex128)
    lib = _lib.load_library()
    p = lambda a: ctypes.c_void_p(a.ctypes.data if a.size else 0)
    check(lib.NUFFT_2D_GPU_host(p(xs), p(ys), p(ss), p(ts), p(fs), p(F), float(eps), xs.size, ss.size),
          "NUFFT_2D_GPU_host")
    return F


def nufft3_params(x: torch.Tensor, y: torch.Tensor, s: torch.Tensor, t: torch.Tensor, eps: float = 1e-6) -> dict:
    """The parameters NUFFT_2D_CPU derives (NFFT3_2D_C.cpp:214-267), from device arrays."""
    dev = x.device
    xs, ys = _vec(x, torch.float64, "x", dev), _vec(y, torch.float64, "y", dev)
    ss, ts = _vec(s, torch.float64, "s", dev), _vec(t, torch.float64, "t", dev)
    out = np.zeros(14)
    lib = _lib.load_library()
    with torch.cuda.device(dev):
        torch.cuda.current_stream(dev).synchronize()
        check(lib.NUFFT_2D_GPU_params(_ptr(xs), _ptr(ys), _ptr(ss), _ptr(ts), float(eps), xs.numel(), ss.numel(),
                                      ctypes.c_void_p(out.ctypes.data)), "NUFFT_2D_GPU_params")
    return {k: (int(v) if k in ("msp", "Mrx", "Mry") else float(v)) for k, v in zip(_KEYS, out)}
