"""Batched type-2 NFFTs of different sizes, over the C ABI (nfft2d_batch_*).

``Batch(sizes, device)`` with ``sizes = [(N1, N2, M), ...]`` builds one plan
per transform (complex64, K = 6, c = 2; N1, N2 even in [2, 1024]: powers of
two from 8 run the Stockham passes, other sizes direct ones).  Every stage
then runs as one kernel launch over all transforms: ``setpts()`` is 5
launches, ``type2()`` and ``type1()`` 4 each, so they can be captured into one
graph.  ``type1`` (N1, N2 >= 22) is the reference's NFFT2_2D_GPU for every
transform, e.g. the far fields of one PSO generation
(``Synthesis/PSO/PSO.cu:348``, ``Synthesis.cu:395``) in one graph.  Arrays are kept by reference (``set_points`` / ``set_io``): keep
the tensors alive while the batch uses them.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check

__all__ = ["Batch"]


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check(t: torch.Tensor, dtype, numel: int, name: str, device) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.device != device:
        raise ValueError(f"{name} must be a tensor on {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.is_conj() or t.is_neg():
        # torch's lazy conjugate / negative views leave the memory unchanged
        raise ValueError(f"{name} is a lazy conjugate or negative view: pass t.resolve_conj().resolve_neg()")
    if t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")


class Batch:
    """Many complex64 type-2 / type-1 NFFTs of different sizes, one launch per stage."""

    def __init__(self, sizes, device=None):
        self.sizes = [(int(a), int(b), int(m)) for a, b, m in sizes]
        self.count = len(self.sizes)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._lib = _lib.load_library()
        n = self.count
        arr = ctypes.c_int * max(n, 1)
        N1 = arr(*[s[0] for s in self.sizes])
        N2 = arr(*[s[1] for s in self.sizes])
        M = arr(*[s[2] for s in self.sizes])
        self._h = ctypes.c_void_p()
        self._keep = ()
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_create(ctypes.byref(self._h), n, N1, N2, M, 0), "nfft2d_batch_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.nfft2d_batch_destroy(h)
            self._h = None

    def _ptrs(self, tensors):
        return (ctypes.c_void_p * max(self.count, 1))(*[ctypes.c_void_p(t.data_ptr()) for t in tensors])

    def set_points(self, xs, ys) -> "Batch":
        if len(xs) != self.count or len(ys) != self.count:
            raise ValueError("one x and one y tensor per transform")
        for t, (x, y) in enumerate(zip(xs, ys)):
            _check(x, torch.float32, self.sizes[t][2], f"x[{t}]", self.device)
            _check(y, torch.float32, self.sizes[t][2], f"y[{t}]", self.device)
        self._pts = (list(xs), list(ys))
        check(self._lib.nfft2d_batch_set_points(self._h, self._ptrs(xs), self._ptrs(ys)), "nfft2d_batch_set_points")
        return self

    def set_io(self, modes, results) -> "Batch":
        if len(modes) != self.count or len(results) != self.count:
            raise ValueError("one mode array and one result array per transform")
        for t, (f, r) in enumerate(zip(modes, results)):
            _check(f, torch.complex64, self.sizes[t][0] * self.sizes[t][1], f"modes[{t}]", self.device)
            _check(r, torch.complex64, self.sizes[t][2], f"results[{t}]", self.device)
        self._io = (list(modes), list(results))
        check(self._lib.nfft2d_batch_set_io(self._h, self._ptrs(modes), self._ptrs(results)), "nfft2d_batch_set_io")
        return self

    def setpts(self) -> None:
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_setpts(self._h, _stream(self.device)), "nfft2d_batch_setpts")

    def type2(self) -> None:
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_type2(self._h, _stream(self.device)), "nfft2d_batch_type2")

    def setpts_type2(self) -> None:
        """setpts() and type2() with the tile sort on the batch's second stream,
        concurrent with the grid passes."""
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_setpts_type2(self._h, _stream(self.device)), "nfft2d_batch_setpts_type2")

    def set_io_type1(self, data, results) -> "Batch":
        """Type-1 arrays: data[t] (M[t] samples), results[t] (N1 x N2 modes); kept by reference."""
        if len(data) != self.count or len(results) != self.count:
            raise ValueError("one sample array and one result array per transform")
        for t, (d, r) in enumerate(zip(data, results)):
            _check(d, torch.complex64, self.sizes[t][2], f"data[{t}]", self.device)
            _check(r, torch.complex64, self.sizes[t][0] * self.sizes[t][1], f"results[{t}]", self.device)
        self._io1 = (list(data), list(results))
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_set_io_type1(self._h, self._ptrs(data), self._ptrs(results)),
                  "nfft2d_batch_set_io_type1")
        return self

    def type1(self) -> None:
        """results[t] = type-1 of data[t] at the sorted points of transform t (NFFT2_2D_GPU per transform)."""
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_type1(self._h, _stream(self.device)), "nfft2d_batch_type1")

    def setpts_type1(self) -> None:
        """setpts() then type1() (9 launches in all)."""
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_setpts_type1(self._h, _stream(self.device)), "nfft2d_batch_setpts_type1")

    def autotune(self, reps: int = 3) -> list:
        """Choose by timing the subproblem size and the direct-gather cut-off per size class and the
        width of the persistent gather (``tuning()``); returns the per-transform subproblem sizes."""
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_batch_autotune(self._h, _stream(self.device), int(reps)), "nfft2d_batch_autotune")
        return self.maxpts()

    def tuning(self) -> dict:
        """The batch's current tuning: per-transform subproblem sizes and direct-gather cut-offs, and the
        number of blocks of the persistent gather."""
        return {"maxpts": self.maxpts(),
                "direct": [self._lib.nfft2d_batch_get_direct(self._h, t) for t in range(self.count)],
                "gather_blocks": self._lib.nfft2d_batch_get_gather_blocks(self._h)}

    def set_direct(self, t: int, npts: int) -> None:
        """Subproblems of transform t with at most npts (0..16) points skip the LDS tile."""
        check(self._lib.nfft2d_batch_set_direct(self._h, int(t), int(npts)), "nfft2d_batch_set_direct")

    def set_gather_blocks(self, blocks: int) -> None:
        check(self._lib.nfft2d_batch_set_gather_blocks(self._h, int(blocks)), "nfft2d_batch_set_gather_blocks")

    def maxpts(self) -> list:
        return [self._lib.nfft2d_batch_get_maxpts(self._h, t) for t in range(self.count)]

    def set_maxpts(self, t: int, maxpts: int) -> None:
        check(self._lib.nfft2d_batch_set_maxpts(self._h, int(t), int(maxpts)), "nfft2d_batch_set_maxpts")
