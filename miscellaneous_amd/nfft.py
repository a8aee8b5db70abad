"""Python mirror of the reference's 2D NFFT interface, over the C ABI.

Reference surface (paths relative to the reference root):

* ``NFFT2_2D_GPU(result, data, x, y, N1, N2, M)`` -- ``NFFT/NFFT2_2D.cu:329,413``:
  type-1, non-uniform data -> N1 x N2 uniform result.
* ``NFFT1_2D_GPU(result, data, x, y, N1, N2, M)`` -- the MATLAB
  ``BP_FBP_FFBP/NFFT1_2D.m``: type-2, N1 x N2 uniform data -> M non-uniform
  results.
* ``NDFT2_2D_GPU(handle, X, Y, u, v, in_, out, Nu, Nv, M, N)`` --
  ``NFFT/NDFT2_2D.cu:37``: direct FP64 sum.
* ``Calculate_cuFFT_plan_C2C_NFFT2_2D`` / ``..._Z2Z_...`` /
  ``Destroy_cuFFT_plan_NFFT2_2D`` -- ``NFFT/NFFT2_2D.cu:49-55`` (here: warm
  or drop the plan cache).

Arguments are torch tensors on a ROCm device (device memory, as in the
reference), complex64 with float32 coordinates or complex128 with float64
coordinates.  Work is enqueued on torch's current stream.  Errors raise
``NFFT2DError`` (the reference aborts the process instead).

``Plan`` is the stream-ordered plan API (create / setpts / execute), the form
used for repeated transforms on the same points.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import NFFT2D_C64, NFFT2D_C128, NFFT2DError, check

__all__ = ["Plan", "NFFT2_2D_GPU", "NFFT1_2D_GPU", "NDFT2_2D_GPU", "ndft_grid_info",
           "Calculate_cuFFT_plan_C2C_NFFT2_2D", "Calculate_cuFFT_plan_Z2Z_NFFT2_2D", "Destroy_cuFFT_plan_NFFT2_2D",
           "NFFT2DError"]

_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}
_CODE = {torch.complex64: NFFT2D_C64, torch.complex128: NFFT2D_C128}


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _require(t: torch.Tensor, dtype, numel: int, name: str, device=None) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be device memory (a ROCm tensor), got {t.device}")
    if device is not None and t.device != device:
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


class Plan:
    """A 2D NFFT plan of size N1 x N2 (c = 2, K = 6) for complex64 or complex128.

    Holds the twiddle / deconvolution tables, the oversampled grid and the
    tile-sorted points; ``setpts`` sorts a point set, after which ``type2``,
    ``type1``, ``adjoint``, ``type2_dx`` and ``type2_grad_x`` run on it.
    """

    def __init__(self, N1: int, N2: int, dtype=torch.complex64, K: int = 6, c: int = 2, device=None):
        if dtype not in _CODE:
            raise TypeError("dtype must be torch.complex64 or torch.complex128")
        self.N1, self.N2, self.dtype, self.rdtype = int(N1), int(N2), dtype, _REAL[dtype]
        dev = torch.device("cuda") if device is None else torch.device(device)
        if dev.type != "cuda":
            raise ValueError(f"Plan needs a ROCm device, got {dev}")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self._lib = _lib.load_library()
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_plan_create(ctypes.byref(self._h), self.N1, self.N2, _CODE[dtype], K, c, 0),
                  "nfft2d_plan_create")
        self.M = None
        self._pts_src = None         # (x, y, versions) of the last setpts_if_changed
        self.pts_version = 0
        self.grid_version = 0        # bumped by every call that rewrites the oversampled grid

    # -------------------------------------------------------------- lifetime
    def close(self) -> None:
        if self._h:
            torch.cuda.synchronize(self.device)
            self._lib.nfft2d_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._h:
                self._lib.nfft2d_plan_destroy(self._h)
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -------------------------------------------------------------- points
    def setpts(self, x: torch.Tensor, y: torch.Tensor) -> "Plan":
        M = x.numel()
        _require(x, self.rdtype, M, "x", self.device)
        _require(y, self.rdtype, M, "y", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_setpts(self._h, _ptr(x), _ptr(y), M, _stream(self.device)), "nfft2d_setpts")
        self.M = M
        self.pts_version += 1
        self._pts_src = None
        return self

    def setpts_if_changed(self, x: torch.Tensor, y: torch.Tensor) -> bool:
        """setpts(x, y) unless the plan's points were last set from these very tensors and neither
        has been modified in place since (same objects, same version counters).  Returns True when
        it sorted.  The plan keeps references to x and y, so their memory cannot be reused for
        other coordinates while the comparison stands."""
        src = self._pts_src
        if (src is not None and src[0] is x and src[1] is y and src[2] == x._version and src[3] == y._version
                and self.M == x.numel()):
            return False
        xc = x.detach().resolve_neg().contiguous()
        yc = y.detach().resolve_neg().contiguous()
        self.setpts(xc, yc)
        self._pts_src = (x, y, x._version, y._version)
        return True

    def set_maxpts(self, maxpts: int) -> None:
        check(self._lib.nfft2d_plan_set_maxpts(self._h, int(maxpts)), "nfft2d_plan_set_maxpts")
        self.M = None
        self._pts_src = None

    def _need_points(self):
        if self.M is None:
            raise NFFT2DError(-5, "plan", "call setpts first")

    # -------------------------------------------------------------- transforms
    def type2(self, f: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """y[l] = sum_k f[k] e^{-2 pi i k.x_l/N};  f: (N1, N2) centred modes."""
        self._need_points()
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        if out is None:
            out = torch.empty(self.M, dtype=self.dtype, device=self.device)
        _require(out, self.dtype, self.M, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2(self._h, _ptr(out), _ptr(f), _stream(self.device)), "nfft2d_type2")
            self.grid_version += 1
        return out

    def setpts_type2(self, x: torch.Tensor, y: torch.Tensor, f: torch.Tensor,
                     out: torch.Tensor | None = None) -> torch.Tensor:
        """setpts(x, y) then type2(f): the tile sort runs on the plan's second
        stream, concurrent with the grid passes; the result is ordered on
        torch's current stream."""
        M = x.numel()
        _require(x, self.rdtype, M, "x", self.device)
        _require(y, self.rdtype, M, "y", self.device)
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        if out is None:
            out = torch.empty(M, dtype=self.dtype, device=self.device)
        _require(out, self.dtype, M, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_setpts_type2(self._h, _ptr(out), _ptr(f), _ptr(x), _ptr(y), M,
                                                _stream(self.device)), "nfft2d_setpts_type2")
            self.grid_version += 1
        self.M = M
        self.pts_version += 1
        self._pts_src = None
        return out

    def setpts_type2_if_changed(self, x: torch.Tensor, y: torch.Tensor, f: torch.Tensor,
                                out: torch.Tensor | None = None) -> torch.Tensor:
        """setpts_type2(x, y, f), or only type2(f) when the plan's points were last set from these
        very tensors, unmodified since (the test of setpts_if_changed).  x and y must be contiguous;
        the plan keeps references to them."""
        src = self._pts_src
        if (src is not None and src[0] is x and src[1] is y and src[2] == x._version and src[3] == y._version
                and self.M == x.numel()):
            return self.type2(f, out=out)
        out = self.setpts_type2(x, y, f, out=out)
        self._pts_src = (x, y, x._version, y._version)
        return out

    def type1(self, d: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """F[k] = sum_l d[l] e^{-2 pi i k.x_l/N}  (the reference NFFT2_2D_GPU)."""
        self._need_points()
        _require(d, self.dtype, self.M, "d", self.device)
        if out is None:
            out = torch.empty(self.N1, self.N2, dtype=self.dtype, device=self.device)
        _require(out, self.dtype, self.N1 * self.N2, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type1(self._h, _ptr(out), _ptr(d), _stream(self.device)), "nfft2d_type1")
            self.grid_version += 1
        return out

    def adjoint(self, g: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """A^H g: F[k] = sum_l g[l] e^{+2 pi i k.x_l/N}."""
        self._need_points()
        _require(g, self.dtype, self.M, "g", self.device)
        if out is None:
            out = torch.empty(self.N1, self.N2, dtype=self.dtype, device=self.device)
        _require(out, self.dtype, self.N1 * self.N2, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_adjoint(self._h, _ptr(out), _ptr(g), _stream(self.device)), "nfft2d_adjoint")
            self.grid_version += 1
        return out

    # -------------------------------------------------------------- stages
    def type2_grid(self, f: torch.Tensor) -> None:
        """Stages (a)+(b): pre-compensate, zero-pad and FFT f into the plan's grid."""
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_grid(self._h, _ptr(f), _stream(self.device)), "nfft2d_type2_grid")
            self.grid_version += 1

    def type2_interp(self, out: torch.Tensor) -> torch.Tensor:
        """Stage (c): gather the plan's grid at the plan's points into out (M)."""
        self._need_points()
        _require(out, self.dtype, self.M, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_interp(self._h, _ptr(out), _stream(self.device)), "nfft2d_type2_interp")
        return out

    def type1_spread(self, d: torch.Tensor) -> None:
        """Stage (d): zero the grid and spread the M samples d onto it."""
        self._need_points()
        _require(d, self.dtype, self.M, "d", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type1_spread(self._h, _ptr(d), _stream(self.device)), "nfft2d_type1_spread")
            self.grid_version += 1

    def type1_finish(self, out: torch.Tensor, sign: int = -1) -> torch.Tensor:
        """FFT + crop + deconvolution of the plan's grid into out (N1 x N2); sign +1 = adjoint."""
        _require(out, self.dtype, self.N1 * self.N2, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type1_finish(self._h, _ptr(out), int(sign), _stream(self.device)),
                  "nfft2d_type1_finish")
        return out

    def type2_dx(self, f: torch.Tensor):
        """(y, dy/dx1, dy/dx2) at the plan's points."""
        self._need_points()
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        y = torch.empty(self.M, dtype=self.dtype, device=self.device)
        d1 = torch.empty_like(y)
        d2 = torch.empty_like(y)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_dx(self._h, _ptr(y), _ptr(d1), _ptr(d2), _ptr(f), _stream(self.device)),
                  "nfft2d_type2_dx")
            self.grid_version += 1
        return y, d1, d2

    def type2_grad_x(self, g: torch.Tensor, f: torch.Tensor):
        """(Re(conj g dy/dx1), Re(conj g dy/dx2)), recomputing the grid from f."""
        self._need_points()
        _require(g, self.dtype, self.M, "g", self.device)
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        gx1 = torch.empty(self.M, dtype=self.rdtype, device=self.device)
        gx2 = torch.empty_like(gx1)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_grad_x(self._h, _ptr(gx1), _ptr(gx2), _ptr(g), _ptr(f),
                                                _stream(self.device)), "nfft2d_type2_grad_x")
            self.grid_version += 1
        return gx1, gx2

    def type2_lsq(self, f: torch.Tensor, y_obs: torch.Tensor, r: torch.Tensor | None = None, keep_r: bool = True):
        """Least squares ||A(x) f - y_obs||^2 in one gather: returns r = 2 (A(x) f - y_obs) (the
        adjoint of r is the gradient w.r.t. f) and the coordinate gradient (gx1, gx2).  The plan
        keeps r for ``lsq_adjoint``; with keep_r=False r is not written out (returned as None)."""
        self._need_points()
        _require(f, self.dtype, self.N1 * self.N2, "f", self.device)
        _require(y_obs, self.dtype, self.M, "y_obs", self.device)
        if keep_r:
            if r is None:
                r = torch.empty(self.M, dtype=self.dtype, device=self.device)
            _require(r, self.dtype, self.M, "r", self.device)
        elif r is not None:
            raise ValueError("type2_lsq: r given with keep_r=False")
        gx1 = torch.empty(self.M, dtype=self.rdtype, device=self.device)
        gx2 = torch.empty_like(gx1)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_lsq(self._h, _ptr(r), _ptr(gx1), _ptr(gx2), _ptr(y_obs), _ptr(f),
                                             _stream(self.device)), "nfft2d_type2_lsq")
            self.grid_version += 1
        return r, gx1, gx2

    def lsq_adjoint(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """dL/df of the least-squares loss of the last ``type2_lsq``: the adjoint of the residual the
        plan holds (``adjoint(r)``, bit for bit on grids of at least 44 x 44), spread in the plan's tile order."""
        self._need_points()
        if out is None:
            out = torch.empty(self.N1, self.N2, dtype=self.dtype, device=self.device)
        _require(out, self.dtype, self.N1 * self.N2, "out", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_lsq_adjoint(self._h, _ptr(out), _stream(self.device)), "nfft2d_lsq_adjoint")
            self.grid_version += 1
        return out

    def type2_grad_x_grid(self, g: torch.Tensor):
        """type2_grad_x on the oversampled grid the plan holds: valid while grid_version is the one
        right after the type2 / type2_grid of the same f (no type-1, adjoint or spread since)."""
        self._need_points()
        _require(g, self.dtype, self.M, "g", self.device)
        gx1 = torch.empty(self.M, dtype=self.rdtype, device=self.device)
        gx2 = torch.empty_like(gx1)
        with torch.cuda.device(self.device):
            check(self._lib.nfft2d_type2_grad_x_grid(self._h, _ptr(gx1), _ptr(gx2), _ptr(g), _stream(self.device)),
                  "nfft2d_type2_grad_x_grid")
        return gx1, gx2


# ------------------------------------------------------------------ drop-in calls
def _dropin(name: str, type2: bool, result, data, x, y, N1: int, N2: int, M: int) -> None:
    if data.dtype not in _CODE:
        raise TypeError(f"{name}: data must be complex64 or complex128")
    dt = data.dtype
    dev = data.device
    _require(data, dt, N1 * N2 if type2 else M, "data", dev)
    _require(result, dt, M if type2 else N1 * N2, "result", dev)
    _require(x, _REAL[dt], M, "x", dev)
    _require(y, _REAL[dt], M, "y", dev)
    lib = _lib.load_library()
    sfx = "_c64" if dt == torch.complex64 else "_c128"
    fn = getattr(lib, name + sfx)
    # The drop-ins run on the legacy default stream, as the reference does;
    # order them after torch's current stream.
    with torch.cuda.device(dev):
        if torch.cuda.current_stream(dev).cuda_stream != 0:
            torch.cuda.current_stream(dev).synchronize()
        check(fn(_ptr(result), _ptr(data), _ptr(x), _ptr(y), int(N1), int(N2), int(M)), name)


def NFFT2_2D_GPU(result, data, x, y, N1: int, N2: int, M: int) -> None:
    """Type-1 NFFT, in place into ``result`` (N1*N2 elements), as NFFT/NFFT2_2D.cu:329."""
    _dropin("NFFT2_2D_GPU", False, result, data, x, y, N1, N2, M)


def NFFT1_2D_GPU(result, data, x, y, N1: int, N2: int, M: int) -> None:
    """Type-2 NFFT, in place into ``result`` (M elements), as BP_FBP_FFBP/NFFT1_2D.m."""
    _dropin("NFFT1_2D_GPU", True, result, data, x, y, N1, N2, M)


def NDFT2_2D_GPU(handle, X, Y, u, v, in_, out, Nu: int, Nv: int, M: int, N: int) -> None:
    """Direct FP64 NED sum into ``out`` (N elements), as NFFT/NDFT2_2D.cu:37-72."""
    dev = in_.device
    for t, n, name in ((X, M, "X"), (Y, M, "Y"), (u, N, "u"), (v, N, "v")):
        _require(t, torch.float64, n, name, dev)
    _require(in_, torch.complex128, M, "in", dev)
    _require(out, torch.complex128, N, "out", dev)
    lib = _lib.load_library()
    with torch.cuda.device(dev):
        if torch.cuda.current_stream(dev).cuda_stream != 0:
            torch.cuda.current_stream(dev).synchronize()
        check(lib.NDFT2_2D_GPU(ctypes.c_void_p(handle if isinstance(handle, int) else 0), _ptr(X), _ptr(Y), _ptr(u),
                               _ptr(v), _ptr(in_), _ptr(out), int(Nu), int(Nv), int(M), int(N)), "NDFT2_2D_GPU")


def ndft_grid_info(device=None) -> tuple[bool, int, int, int]:
    """(is_grid, nu, nv, vmajor) of the last ``NDFT2_2D_GPU`` call with outputs on ``device``: whether its
    outputs were detected as a grid of nu x nv values and took the matrix-core contraction.  Synchronises."""
    lib = _lib.load_library()
    vals = [ctypes.c_int(0) for _ in range(4)]
    with torch.cuda.device(device if device is not None else torch.cuda.current_device()):
        check(lib.nfft2d_ndft_grid_info(*[ctypes.byref(x) for x in vals]), "nfft2d_ndft_grid_info")
    return bool(vals[0].value), vals[1].value, vals[2].value, vals[3].value


def Calculate_cuFFT_plan_C2C_NFFT2_2D(N1: int, N2: int) -> None:
    _lib.load_library().Calculate_cuFFT_plan_C2C_NFFT2_2D(int(N1), int(N2))


def Calculate_cuFFT_plan_Z2Z_NFFT2_2D(N1: int, N2: int) -> None:
    _lib.load_library().Calculate_cuFFT_plan_Z2Z_NFFT2_2D(int(N1), int(N2))


def Destroy_cuFFT_plan_NFFT2_2D() -> None:
    _lib.load_library().Destroy_cuFFT_plan_NFFT2_2D()
