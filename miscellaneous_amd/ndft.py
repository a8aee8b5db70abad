"""Type-2 NDFT as a split-precision MFMA contraction, over the C ABI.

``NdftPlan(N1, N2, precision)`` evaluates the direct sum

    y[m] = sum_{k1,k2} f[k1, k2] e^{-2 pi i (k1 x[m] / N1 + k2 y[m] / N2)}

(the sum of ``NFFT/NDFT2_2D.cu:15-32,67`` and ``NFFT/NFFT2/Matlab/NDFT2_2D.m``,
taken over the equispaced modes) as ``rowsum(A o (B f^T))`` on the matrix
cores.  ``precision`` is ``"fp32"``, ``"bf16x3"`` (split BF16, three products)
or ``"fp16"`` (per-row power-of-two scaling); accumulation is FP32 in all
three.  Modes are centred, row-major ``[k1][k2]``; coordinates are float32 grid
units, as for ``NFFT1_2D_GPU``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import NFFT2D_NDFT_BF16X3, NFFT2D_NDFT_FP16, NFFT2D_NDFT_FP32, check

PRECISIONS = {"fp32": NFFT2D_NDFT_FP32, "bf16x3": NFFT2D_NDFT_BF16X3, "fp16": NFFT2D_NDFT_FP16}

__all__ = ["NdftPlan", "PRECISIONS"]


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_tensor(t: torch.Tensor, dtype, numel: int, name: str, device) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a ROCm tensor")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.is_conj() or t.is_neg():
        # torch's lazy conjugate / negative views leave the memory unchanged
        raise ValueError(f"{name} is a lazy conjugate or negative view: pass t.resolve_conj().resolve_neg()")
    if t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")


class NdftPlan:
    """Plan of the type-2 NDFT contraction for N1 x N2 modes in one precision."""

    def __init__(self, N1: int, N2: int, precision: str = "bf16x3", device=None):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
        self.N1, self.N2, self.precision = int(N1), int(N2), precision
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._lib = _lib.load_library()
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_ndft_plan_create(ctypes.byref(self._h), self.N1, self.N2, PRECISIONS[precision]),
                  "nfft2d_ndft_plan_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.nfft2d_ndft_plan_destroy(h)
            self._h = None

    def set_modes(self, f: torch.Tensor) -> "NdftPlan":
        """Prepare the mode operand from f (N1 x N2 complex64)."""
        _check_tensor(f, torch.complex64, self.N1 * self.N2, "f", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_ndft_set_modes(self._h, ctypes.c_void_p(f.data_ptr()), _stream(self.device)),
                  "nfft2d_ndft_set_modes")
        return self

    def eval(self, x: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Contract the prepared modes with the points (x, y): M complex64 results."""
        M = x.numel()
        _check_tensor(x, torch.float32, M, "x", self.device)
        _check_tensor(y, torch.float32, M, "y", self.device)
        if out is None:
            out = torch.empty(M, dtype=torch.complex64, device=self.device)
        _check_tensor(out, torch.complex64, M, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_ndft_eval(self._h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr()), int(M), _stream(self.device)),
                  "nfft2d_ndft_eval")
        return out

    def type2(self, f: torch.Tensor, x: torch.Tensor, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """set_modes(f) then eval(x, y)."""
        return self.set_modes(f).eval(x, y, out)
