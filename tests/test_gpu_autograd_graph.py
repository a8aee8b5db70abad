"""The torch.autograd form of the gradient (forward, loss, backward to f and x) captured whole
into one HIP graph and replayed: every replay, with new modes and observations written into the
captured inputs between replays, must give what the same step gives eagerly on the same plan.

The points stay fixed across replays: the forward's check that skips the tile sort for unchanged
points runs on the host, at capture time.
"""
import numpy as np
import pytest
import torch

import miscellaneous_amd as m

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def cn(gen, shape):
    return torch.complex(torch.randn(shape, generator=gen), torch.randn(shape, generator=gen)).div(2 ** 0.5).to(DEV)


def rel(a, b):
    return ((a - b).norm() / b.norm()).item()


def test_autograd_step_graph_replays_match_eager():
    N, M = 64, 20000
    gen = torch.Generator().manual_seed(31)
    x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(DEV).requires_grad_(True)
    x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(DEV).requires_grad_(True)
    f = cn(gen, (N, N)).requires_grad_(True)
    y_obs = cn(gen, (M,))
    plan = m.Plan(N, N, torch.complex64, device=DEV)

    def step():
        loss = ((m.nfft2d_type2(f, x1, x2, plan) - y_obs).abs() ** 2).sum()
        loss.backward()
        return loss

    side = torch.cuda.Stream(device=DEV)
    side.wait_stream(torch.cuda.current_stream(DEV))
    with torch.cuda.stream(side):          # warm-up: sort, spread tiles, autograd state
        for _ in range(3):
            f.grad = x1.grad = x2.grad = None
            step()
    torch.cuda.current_stream(DEV).wait_stream(side)
    torch.cuda.synchronize()
    f.grad = x1.grad = x2.grad = None
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        static_loss = step()
    gf, gx1, gx2 = f.grad, x1.grad, x2.grad
    for rep in range(4):
        print(f"case autograd graph replay {rep}", flush=True)
        with torch.no_grad():
            f.copy_(cn(gen, (N, N)))
            y_obs.copy_(cn(gen, (M,)))
        g.replay()
        torch.cuda.synchronize()
        got = [static_loss.clone(), gf.clone(), gx1.clone(), gx2.clone()]
        # eager on the same plan (same sorted points)
        f.grad = x1.grad = x2.grad = None
        loss = step()
        torch.cuda.synchronize()
        want = [loss.detach(), f.grad, x1.grad, x2.grad]
        for name, a, b in zip(("loss", "grad f", "grad x1", "grad x2"), got, want):
            e = rel(a, b)
            assert e <= 1e-6, f"replay {rep}: {name} rel {e:.2e}"
        f.grad, x1.grad, x2.grad = gf, gx1, gx2   # the graph's own gradient tensors again
