"""The far-field step of one PSO generation of the reflectarray synthesis, batched.

Reference (paths relative to the reference root): for every particle,
``CostFunctionalCalculation`` (``Synthesis/PSO/PSO.cu:342-363``) calls
``raCostFunctional``, whose ``raFarFieldCalculation`` (``Synthesis/PSO/Synthesis.cu:360-420``)
scales the element positions to grid units (``:389-390``), calls
``NFFT2_2D_GPU(far_field, excitations, X_scaled, Y_scaled, 2 Nv, 2 Nu, M_x M_y)``
(``:395``) on a (2 Nv) x (2 Nu) grid and multiplies the far field by a filter
(``farFieldFiltering``, ``:345-351``): one transform, one set of allocations and
one synchronising call per particle, 250 particles per generation.

Here the whole generation's far fields are ONE batch: the tile sort, the
spread, the FFT passes and the filter of every particle in 10 kernel launches,
captured into one graph and replayed every generation with the new positions
and excitations written into its input arrays.  The aperture field, the
positions (Zernike / Lagrange expansions) and the cost functional around this
step are the synthesis code, not the NFFT, and are not part of this library.
"""
from __future__ import annotations

import torch

from .batch import Batch

__all__ = ["FarFieldGeneration"]


class FarFieldGeneration:
    """Far fields of `particles` particles, each M elements, on (2 Nv) x (2 Nu) grids.

    Inputs live in the tensors ``excitations`` (particles, M) complex64 and ``X``, ``Y``
    (particles, M) float32 (grid units, i.e. already scaled as Synthesis.cu:389-390 does),
    and ``filter`` ((2 Nv) (2 Nu),) float32; ``far_field`` (particles, 2 Nv, 2 Nu) complex64
    is the output.  Write new values into the inputs in place between calls.
    """

    def __init__(self, particles: int, Nu: int, Nv: int, M: int, device=None):
        dev = torch.device("cuda") if device is None else torch.device(device)
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.P, self.Nu, self.Nv, self.M = int(particles), int(Nu), int(Nv), int(M)
        N1, N2 = 2 * self.Nv, 2 * self.Nu            # NFFT2_2D_GPU(..., 2 * Nv, 2 * Nu, ...)
        self.excitations = torch.zeros(self.P, self.M, dtype=torch.complex64, device=self.device)
        self.X = torch.zeros(self.P, self.M, dtype=torch.float32, device=self.device)
        self.Y = torch.zeros(self.P, self.M, dtype=torch.float32, device=self.device)
        self.filter = torch.ones(N1 * N2, dtype=torch.float32, device=self.device)
        self.far_field = torch.zeros(self.P, N1, N2, dtype=torch.complex64, device=self.device)
        self.batch = Batch([(N1, N2, self.M)] * self.P, device=self.device)
        self.batch.set_points(list(self.X), list(self.Y))
        self.batch.set_io_type1(list(self.excitations), [f.view(-1) for f in self.far_field])
        self._graph = None

    def run(self) -> torch.Tensor:
        """One generation, eagerly: batched setpts + type-1, then the filter."""
        self.batch.setpts_type1()
        self.far_field.mul_(self.filter.view(1, *self.far_field.shape[1:]))
        return self.far_field

    def capture(self) -> "torch.cuda.CUDAGraph":
        """Capture run() into a graph (after one eager call that sizes everything)."""
        self.run()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize(self.device)
        self._graph = g
        return g

    def replay(self) -> torch.Tensor:
        if self._graph is None:
            self.capture()
        self._graph.replay()
        return self.far_field
