"""Batched 1-D NFFTs over the C ABI (nfft1d_*), complex64, K = 6, c = 2.

Reference (paths relative to the reference root):

* ``NFFT1_1D(data, x)`` -- ``BP_FBP_FFBP/NFFT1_1D.m:1``, the type-2 the
  backprojection calls once per pulse (``BP_FBP_FFBP/BP_NFFT.m:18``,
  ``BP_POLAR_NFFT.m:19``): N modes to the M points x.
* ``NFFT2_1D(data, x, N, c, K)`` -- ``NFFT/NFFT2/1D/Matlab/NFFT2_1D.m:1``, the
  type-1 of the 1-D aperture synthesis (``Synthesis/Aper_Synth_1D/Aper_Synth_1D.m:74``):
  M samples at x to N modes.

Both take one line (1-D tensors) or a batch of lines (2-D tensors, one line
per row) and run every line of the batch in one launch per stage
(``Plan1D``).  Coordinates are in grid units (period N); modes are centred
(k + N/2).  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check

__all__ = ["Plan1D", "NFFT1_1D", "NFFT2_1D"]


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _as_lines(t: torch.Tensor, dtype, name: str, device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.device != device:
        raise ValueError(f"{name} must be a tensor on {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if t.is_conj() or t.is_neg():
        raise ValueError(f"{name} is a lazy conjugate or negative view: pass t.resolve_conj().resolve_neg()")
    if t.dim() not in (1, 2):
        raise ValueError(f"{name} must be one line (1-D) or a batch of lines (2-D)")
    return t.reshape(1, -1).contiguous() if t.dim() == 1 else t.contiguous()


class Plan1D:
    """Batched 1-D NFFTs of N modes (complex64): the twiddle and deconvolution
    tables, and a grid that grows to the largest batch seen."""

    def __init__(self, N: int, device=None):
        self.N = int(N)
        dev = torch.device("cuda") if device is None else torch.device(device)
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self._lib = _lib.load_library()
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.nfft1d_plan_create(ctypes.byref(self._h), self.N), "nfft1d_plan_create")

    def __del__(self):
        try:
            if self._h:
                self._lib.nfft1d_plan_destroy(self._h)
        except Exception:
            pass

    def type2(self, data: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """result[b, l] = sum_k data[b, k + N/2] e^{-2 pi i k x[b, l] / N}: data (B, N), x (B, M) -> (B, M)."""
        d = _as_lines(data, torch.complex64, "data", self.device)
        xs = _as_lines(x, torch.float32, "x", self.device)
        B, M = xs.shape
        if d.shape != (B, self.N):
            raise ValueError(f"data must be ({B}, {self.N}) for x of shape {tuple(xs.shape)}, got {tuple(d.shape)}")
        out = torch.empty(B, M, dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft1d_type2(self._h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(d.data_ptr()),
                                         ctypes.c_void_p(xs.data_ptr()), M, B, _stream(self.device)), "nfft1d_type2")
        return out if x.dim() == 2 else out[0]

    def type1(self, data: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """result[b, k + N/2] = sum_l data[b, l] e^{-2 pi i k x[b, l] / N}: data, x (B, M) -> (B, N)."""
        d = _as_lines(data, torch.complex64, "data", self.device)
        xs = _as_lines(x, torch.float32, "x", self.device)
        if d.shape != xs.shape:
            raise ValueError(f"data {tuple(d.shape)} and x {tuple(xs.shape)} must have one shape")
        B, M = xs.shape
        out = torch.empty(B, self.N, dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft1d_type1(self._h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(d.data_ptr()),
                                         ctypes.c_void_p(xs.data_ptr()), M, B, _stream(self.device)), "nfft1d_type1")
        return out if x.dim() == 2 else out[0]


_plans: dict = {}


def _plan(N: int, device: torch.device) -> Plan1D:
    key = (N, device)
    p = _plans.get(key)
    if p is None:
        if len(_plans) >= 8:
            _plans.clear()
        p = _plans[key] = Plan1D(N, device)
    return p


def NFFT1_1D(data: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """BP_FBP_FFBP/NFFT1_1D.m: the N modes in ``data`` (last dimension) at the points ``x``."""
    return _plan(int(data.shape[-1]), data.device).type2(data, x)


def NFFT2_1D(data: torch.Tensor, x: torch.Tensor, N: int, c: int = 2, K: int = 6) -> torch.Tensor:
    """NFFT/NFFT2/1D/Matlab/NFFT2_1D.m (``NFFT2_1D(data, x, M, c, K)`` there, M the output size):
    the samples ``data`` at ``x`` to N centred modes.  Only c = 2, K = 6 (the reference's values)."""
    if c != 2 or K != 6:
        raise ValueError("only c = 2 and K = 6 are supported")
    return _plan(int(N), data.device).type1(data, x)
