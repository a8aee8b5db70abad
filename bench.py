#!/usr/bin/env python3
"""Benchmark of the 2D NFFT hot path on one MI355X.

Headline metric: 2D type-2 NFFT throughput in Mpts/s on the polar
backprojection workload -- 512 x 512 complex64 CN(0,1) modes, M = 1,048,576
points on 512 equispaced angles in [0, pi) x 2048 radii in [-1/2, 1/2), K = 6,
c = 2 -- with its rel-L2 against an FP64 direct sum (4096 seeded points).

One step = setpts (tile sort of the M points) + type-2 (window
pre-compensation + zero padding + 2-D FFT + (2K+1)^2 gather), replayed from a
HIP graph so that host launch overhead is not measured.  The exec-only rate
(points already sorted) and the per-kernel times (HIP events on the stream the
kernels run on) are reported beside it.

Secondary: fwd+adjoint gradient steps/s on 256 x 256 modes, M = 262,144
uniform points (loss ||A(x) f - y||^2, gradient w.r.t. f and x, gradient
descent update), also graph-replayed.

Also reported (skip with --no-extra): the batch of 256 uneven type-2 NFFTs
(batched setpts + type-2 in one graph, against the same transforms one plan at
a time), the type-2 NDFT as an MFMA contraction in FP32 / BF16x3 / FP16 with
its MFMA utilisation, the 1024^2 / M=4M uniform type-2, the type-1 and the
complex128 forms on the polar set, and the band-limited polar interpolation
(knabNFFT2D) built on the type-2.

cpu_baseline: the CPU restatement of the reference algorithm (FP32,
OpenMP over the host cores) on the same type-2 workload, timed for a bounded
number of repetitions.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def cn01(gen, shape, device):
    re = torch.randn(shape, generator=gen, dtype=torch.float32, device="cpu")
    im = torch.randn(shape, generator=gen, dtype=torch.float32, device="cpu")
    return (torch.complex(re, im) / np.sqrt(2.0)).to(device)


def polar_points(N: int, n_ang: int = 512, n_rad: int = 2048):
    th = np.arange(n_ang) * np.pi / n_ang
    r = -0.5 + np.arange(n_rad) / n_rad
    R, TH = np.meshgrid(r, th)
    x1 = (R * np.cos(TH)).astype(np.float32).ravel() * np.float32(N)
    x2 = (R * np.sin(TH)).astype(np.float32).ravel() * np.float32(N)
    return x1.astype(np.float32), x2.astype(np.float32)


def direct_type2_subset(f: torch.Tensor, x1: torch.Tensor, x2: torch.Tensor) -> torch.Tensor:
    """FP64 direct sum y_l = sum_k f_k e^{-2 pi i (k1 x1_l/N1 + k2 x2_l/N2)} in torch (separable)."""
    N1, N2 = f.shape
    k1 = torch.arange(-N1 // 2, N1 // 2, device=f.device, dtype=torch.float64)
    k2 = torch.arange(-N2 // 2, N2 // 2, device=f.device, dtype=torch.float64)
    ph1 = torch.outer(x1.double(), k1) / N1
    ph2 = torch.outer(x2.double(), k2) / N2
    A = torch.polar(torch.ones_like(ph1), -2 * np.pi * (ph1 - torch.floor(ph1)))
    B = torch.polar(torch.ones_like(ph2), -2 * np.pi * (ph2 - torch.floor(ph2)))
    G = B @ f.to(torch.complex128).T.contiguous()
    return (A * G).sum(dim=1)


def capture(fn, warm: int = 3):
    """Capture fn() into a CUDA/HIP graph after a few eager warm-up calls."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warm):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def time_graph(g, steps: int, warmup: int) -> float:
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def stage_ms(fn, n):
    """Per-stage launch duration: HIP events on the stream the kernels run on (torch's current
    stream, which the library launches on), around n back-to-back launches of the stage (the queue
    stays full, so the average is the stage's device time per launch)."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def stage_roofline(nbytes: int, ms: float) -> dict:
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"algorithmic_bytes": nbytes, "us": ms * 1e3, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def bench_type2(args, dev):
    import miscellaneous_amd as m

    N = 512
    x1n, x2n = polar_points(N)
    M = len(x1n)
    gen = torch.Generator().manual_seed(SEED)
    f = cn01(gen, (N, N), dev)
    x1 = torch.from_numpy(x1n).to(dev)
    x2 = torch.from_numpy(x2n).to(dev)
    y = torch.empty(M, dtype=torch.complex64, device=dev)
    plan = m.Plan(N, N, torch.complex64, device=dev)
    plan.setpts(x1, x2)
    plan.type2(f, out=y)
    torch.cuda.synchronize()

    # accuracy on 4096 seeded points against an FP64 direct sum
    sel = torch.from_numpy(np.random.default_rng(SEED).choice(M, 4096, replace=False)).to(dev)
    y_ref = direct_type2_subset(f, x1[sel], x2[sel])
    rel = ((y[sel].to(torch.complex128) - y_ref).norm() / y_ref.norm()).item()

    n_ev = max(args.steps, 20)
    setpts_ms = stage_ms(lambda: plan.setpts(x1, x2), n_ev)
    grid_ms = stage_ms(lambda: plan.type2_grid(f), n_ev)
    gather_ms = stage_ms(lambda: plan.type2_interp(y), n_ev)

    def step_full():        # tile sort on the plan's second stream, beside the grid passes
        plan.setpts_type2(x1, x2, f, out=y)

    def step_serial():      # the same work with every stage on one stream
        plan.setpts(x1, x2)
        plan.type2(f, out=y)

    def step_exec():
        plan.type2(f, out=y)

    g_full = capture(step_full)
    g_serial = capture(step_serial)
    g_exec = capture(step_exec)
    s_full = time_graph(g_full, args.steps, args.warmup)
    s_serial = time_graph(g_serial, args.steps, args.warmup)
    s_exec = time_graph(g_exec, args.steps, args.warmup)

    # the timed step's result is the eager result
    ref = y.clone()
    g_full.replay()
    torch.cuda.synchronize()
    assert torch.equal(ref, y), "graph replay of setpts_type2 differs from the eager type-2"

    # graph replays must reproduce the eager result
    ref = y.clone()
    g_exec.replay()
    torch.cuda.synchronize()
    assert torch.equal(ref, y), "graph replay differs from the eager type-2"

    gather_bytes = 32 * N * N + 16 * M          # grid read once + coordinates + output
    type2_bytes = 200 * N * N + 16 * M          # stages (a)+(b)+(c)
    return dict(N=N, M=M, rel_l2=rel, s_full=s_full, s_serial=s_serial, s_exec=s_exec, setpts_ms=setpts_ms, grid_ms=grid_ms,
                gather_ms=gather_ms, gather_bytes=gather_bytes, type2_bytes=type2_bytes, plan=plan, f=f, x1n=x1n,
                x2n=x2n)


def bench_grad(args, dev):
    import miscellaneous_amd as m

    N, M = 256, 262144
    gen = torch.Generator().manual_seed(SEED + 1)
    x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    f_true = cn01(gen, (N, N), dev)
    plan = m.Plan(N, N, torch.complex64, device=dev)
    plan.setpts(x1, x2)
    y_obs = plan.type2(f_true) + cn01(gen, (M,), dev) * np.sqrt(1e-3)
    f = torch.zeros(N, N, dtype=torch.complex64, device=dev)
    gf = torch.empty(N, N, dtype=torch.complex64, device=dev)
    # GD is stable for lr < 1 / lambda_max(A^H A); for M/N^2 = 4 uniform points
    # lambda_max is about M (1 + sqrt(N^2/M))^2 = 2.25 M
    lr = 1.0 / (4 * M)

    def loss_now() -> float:                       # eager, outside the graph
        return float(((plan.type2(f) - y_obs).abs() ** 2).sum().item())

    def step():
        # forward, residual r = 2 (y - y_obs) (kept in the plan) and dL/dx in one gather
        plan.type2_lsq(f, y_obs, keep_r=False)
        plan.lsq_adjoint(out=gf)                   # dL/df = A^H r
        f.add_(gf, alpha=-lr)

    g_step = capture(step)
    f.zero_()
    torch.cuda.synchronize()
    loss0 = loss_now()
    n = max(args.steps, 10)
    s = time_graph(g_step, n, 2)
    total_steps = n + 2
    while total_steps < 200:
        g_step.replay()
        total_steps += 1
    torch.cuda.synchronize()
    return dict(N=N, M=M, s=s, loss0=loss0, loss_final=loss_now(), steps=total_steps, lr=lr,
                bytes=408 * N * N + 80 * M)


def batch_sizes(seed: int, count: int = 256):
    """The batch workload: N uniform in {64, 128, 256, 512} (square), M log-uniform in [4096, 262144]."""
    rng = np.random.default_rng(seed)
    Ns = rng.choice([64, 128, 256, 512], size=count)
    Ms = np.round(np.exp(rng.uniform(np.log(4096), np.log(262144), size=count))).astype(int)
    return [(int(n), int(n), int(mm)) for n, mm in zip(Ns, Ms)], rng


def batch_points(rng, N, M):
    """Half uniform in [-1/2, 1/2)^2, half Gaussian (sigma 0.05) wrapped to [-1/2, 1/2); grid units."""
    h = M // 2
    u = rng.random((h, 2)) - 0.5
    g = (rng.normal(0.0, 0.05, size=(M - h, 2)) + 0.5) % 1.0 - 0.5
    p = np.concatenate([u, g]).astype(np.float32)
    return (p[:, 0] * np.float32(N)).astype(np.float32), (p[:, 1] * np.float32(N)).astype(np.float32)


def bench_batch(args, dev):
    """256 type-2 NFFTs of uneven sizes: one graph replay of the batched setpts + type-2 (9 launches),
    against the same transforms run one plan at a time (also graph-replayed)."""
    import miscellaneous_amd as m

    sizes, rng = batch_sizes(SEED + 3)
    gen = torch.Generator().manual_seed(SEED + 3)
    fs = [cn01(gen, (n1, n2), dev) for (n1, n2, _) in sizes]
    pts = [batch_points(rng, n1, mm) for (n1, _, mm) in sizes]
    xs = [torch.from_numpy(a).to(dev) for a, _ in pts]
    ys = [torch.from_numpy(b).to(dev) for _, b in pts]
    outs = [torch.empty(mm, dtype=torch.complex64, device=dev) for (_, _, mm) in sizes]
    b = m.Batch(sizes, device=dev).set_points(xs, ys).set_io(fs, outs)

    def step():             # tile sort on the batch's second stream, beside the grid passes
        b.setpts_type2()

    g0 = capture(step)      # the defaults, before tuning: subproblems of 256, direct cut-off 8, 1024 blocks
    s_default = time_graph(g0, max(args.steps // 4, 10), 3)
    del g0
    t0 = time.perf_counter()
    choice = b.autotune(reps=5)
    tune_s = time.perf_counter() - t0
    tune = b.tuning()

    def step_serial():
        b.setpts()
        b.type2()

    g = capture(step)
    s_batch = time_graph(g, max(args.steps // 4, 10), 3)
    g_serial = capture(step_serial)
    s_serial = time_graph(g_serial, max(args.steps // 4, 10), 3)
    del g_serial

    # accuracy: 256 seeded outputs of 16 transforms against an FP64 direct sum
    step()
    torch.cuda.synchronize()
    worst = 0.0
    sel_rng = np.random.default_rng(SEED + 4)
    for t in sel_rng.choice(len(sizes), 16, replace=False):
        mm = sizes[t][2]
        sel = torch.from_numpy(sel_rng.choice(mm, min(mm, 256), replace=False)).to(dev)
        ref = direct_type2_subset(fs[t], xs[t][sel], ys[t][sel])
        worst = max(worst, ((outs[t][sel].to(torch.complex128) - ref).norm() / ref.norm()).item())

    # the same transforms through one plan each (the single-transform FP32 path)
    plans = [m.Plan(n1, n2, torch.complex64, device=dev) for (n1, n2, _) in sizes]
    outs2 = [torch.empty_like(o) for o in outs]

    def step_plain():
        for p, f, x, y, o in zip(plans, fs, xs, ys, outs2):
            p.setpts(x, y)
            p.type2(f, out=o)

    g2 = capture(step_plain)
    s_plain = time_graph(g2, max(args.steps // 20, 3), 1)
    same = all(torch.equal(a, c) for a, c in zip(outs, outs2))
    total_m = sum(mm for (_, _, mm) in sizes)
    nbytes = sum(200 * n1 * n2 + 16 * mm for (n1, n2, mm) in sizes)
    tuned, direct = {}, {}
    for (n1, _, _), c, d in zip(sizes, choice, tune["direct"]):
        tuned[str(n1)] = c
        direct[str(n1)] = d
    del plans, g2
    return {"config": "256 type-2 NFFTs, N uniform in {64,128,256,512} (square), M log-uniform in [4096, 262144], "
                      "half uniform / half Gaussian (sigma 0.05, wrapped) points, complex64, K=6, c=2; "
                      "step = batched setpts + type-2 (9 launches, the tile sort beside the grid passes) replayed from one graph",
            "transforms": len(sizes), "total_points": total_m, "value": total_m / s_batch / 1e6, "unit": "Mpts/s",
            "ms_per_step": s_batch * 1e3, "one_stream_ms_per_step": s_serial * 1e3, "plain_ms_per_step": s_plain * 1e3, "speedup_vs_plain": s_plain / s_batch,
            "identical_to_plain": same, "rel_l2_max_16x256": worst,
            "hbm_frac": nbytes / s_batch / (HBM_PEAK_GBS * 1e9), "algorithmic_bytes": nbytes,
            "autotuned_maxpts_by_N": tuned, "autotuned_direct_cutoff_by_N": direct,
            "autotuned_gather_blocks": tune["gather_blocks"], "default_tuning_ms_per_step": s_default * 1e3,
            "autotune_s": tune_s}


def bench_ndft(args, dev):
    """Type-2 NDFT as an MFMA contraction, 256 x 256 modes, M = 65536 uniform points, three precisions."""
    import miscellaneous_amd as m

    N, M = 256, 65536
    gen = torch.Generator().manual_seed(SEED + 5)
    f = cn01(gen, (N, N), dev)
    x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    sel = torch.from_numpy(np.random.default_rng(SEED + 5).choice(M, 4096, replace=False)).to(dev)
    ref = direct_type2_subset(f, x1[sel], x2[sel])
    peak = {"fp32": 157.3e12, "bf16x3": 2.5e15, "fp16": 2.5e15}
    prods = {"fp32": 1, "bf16x3": 3, "fp16": 1}
    res = {}
    for prec in ("fp32", "bf16x3", "fp16"):
        p = m.NdftPlan(N, N, prec, device=dev).set_modes(f)
        out = torch.empty(M, dtype=torch.complex64, device=dev)
        p.eval(x1, x2, out)
        torch.cuda.synchronize()
        rel = ((out[sel].to(torch.complex128) - ref).norm() / ref.norm()).item()
        reps = 5 if prec == "fp32" else 20
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            p.eval(x1, x2, out)
        ev[1].record()
        torch.cuda.synchronize()
        t = ev[0].elapsed_time(ev[1]) / reps * 1e-3
        flops = 8.0 * M * N * N * prods[prec]
        res[prec] = {"ms": t * 1e3, "rel_l2": rel, "executed_tflops": flops / t / 1e12,
                     "mfma_util": flops / t / peak[prec], "useful_tflops": 8.0 * M * N * N / t / 1e12}
    return {"config": "type-2 NDFT, 256x256 CN(0,1) modes, M=65536 uniform points, T = B F^T on MFMA "
                      "(k_ndft_t2), rel-L2 on 4096 points vs FP64 direct sum; MFMA peaks: FP32 157.3, "
                      "BF16/FP16 2500 TFLOP/s (spec)", **res}


def bench_large(args, dev):
    """Type-2 NFFT at 1024 x 1024 modes, M = 4194304 uniform points (setpts + type-2, graph-replayed)."""
    import miscellaneous_amd as m

    N, M = 1024, 4194304
    gen = torch.Generator().manual_seed(SEED + 6)
    f = cn01(gen, (N, N), dev)
    x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
    y = torch.empty(M, dtype=torch.complex64, device=dev)
    plan = m.Plan(N, N, torch.complex64, device=dev)

    def step():
        plan.setpts(x1, x2)
        plan.type2(f, out=y)

    g = capture(step)
    s = time_graph(g, max(args.steps // 10, 5), 2)
    sel = torch.from_numpy(np.random.default_rng(SEED + 6).choice(M, 4096, replace=False)).to(dev)
    ref = direct_type2_subset(f, x1[sel], x2[sel])
    rel = ((y[sel].to(torch.complex128) - ref).norm() / ref.norm()).item()
    return {"config": "type-2 NFFT 1024x1024 modes, M=4194304 uniform, K=6, c=2; step = setpts + type-2",
            "value": M / s / 1e6, "unit": "Mpts/s", "ms_per_step": s * 1e3, "rel_l2": rel,
            "hbm_frac": (200 * N * N + 16 * M) / s / (HBM_PEAK_GBS * 1e9)}


def direct_type1_subset(d: torch.Tensor, x1: torch.Tensor, x2: torch.Tensor, k1: torch.Tensor, k2: torch.Tensor,
                        N1: int, N2: int) -> torch.Tensor:
    """FP64 direct sum F_s = sum_l d_l e^{-2 pi i (k1_s x1_l/N1 + k2_s x2_l/N2)} for selected modes s, in point chunks."""
    out = torch.zeros(k1.numel(), dtype=torch.complex128, device=d.device)
    for a in range(0, d.numel(), 65536):
        ph = (torch.outer(x1[a:a + 65536].double(), k1.double()) / N1 +
              torch.outer(x2[a:a + 65536].double(), k2.double()) / N2)
        E = torch.polar(torch.ones_like(ph), -2 * np.pi * (ph - torch.floor(ph)))
        out += (d[a:a + 65536].to(torch.complex128)[:, None] * E).sum(dim=0)
    return out


def bench_type1(args, dev):
    """Type-1 (the reference's NFFT2_2D_GPU direction) on the polar set: 512 x 512 modes, M = 1048576,
    step = setpts + type-1 (spread, FFT, crop + deconvolution), graph-replayed; and the type-1 alone."""
    import miscellaneous_amd as m

    N = 512
    x1n, x2n = polar_points(N)
    M = len(x1n)
    gen = torch.Generator().manual_seed(SEED + 7)
    d = cn01(gen, (M,), dev)
    x1 = torch.from_numpy(x1n).to(dev)
    x2 = torch.from_numpy(x2n).to(dev)
    F = torch.empty(N, N, dtype=torch.complex64, device=dev)
    plan = m.Plan(N, N, torch.complex64, device=dev)

    def step():
        plan.setpts(x1, x2)
        plan.type1(d, out=F)

    def step_exec():
        plan.type1(d, out=F)

    g = capture(step)
    s = time_graph(g, max(args.steps // 4, 10), 3)
    g2 = capture(step_exec)
    s2 = time_graph(g2, max(args.steps // 4, 10), 3)
    n_ev = max(args.steps, 20)
    spread_ms = stage_ms(lambda: plan.type1_spread(d), n_ev)
    finish_ms = stage_ms(lambda: plan.type1_finish(F), n_ev)
    rng = np.random.default_rng(SEED + 7)
    sel = torch.from_numpy(rng.choice(N * N, 256, replace=False)).to(dev)
    k1 = sel // N - N // 2
    k2 = sel % N - N // 2
    ref = direct_type1_subset(d, x1, x2, k1, k2, N, N)
    got = F.reshape(-1)[sel].to(torch.complex128)
    rel = ((got - ref).norm() / ref.norm()).item()
    nbytes = 176 * N * N + 16 * M
    return {"config": "type-1 NFFT (NFFT2_2D_GPU direction), 512x512 modes, M=1048576 polar points, K=6, c=2; "
                      "step = setpts + type-1",
            "value": M / s / 1e6, "unit": "Mpts/s", "ms_per_step": s * 1e3,
            "exec_only_ms": s2 * 1e3, "rel_l2_256_modes": rel,
            "stage_roofline": {"d_spread": stage_roofline(32 * N * N + 16 * M, spread_ms),
                               "fft_crop_deconvolve": stage_roofline(144 * N * N, finish_ms)},
            "hbm_frac_exec": nbytes / s2 / (HBM_PEAK_GBS * 1e9)}


def bench_fp64(args, dev):
    """The complex128 paths on the polar set (512 x 512, M = 1048576), graph-replayed: setpts + type-2,
    and setpts + type-1 (the reference's double NFFT2_2D_GPU)."""
    import miscellaneous_amd as m

    N = 512
    x1n, x2n = polar_points(N)
    M = len(x1n)
    gen = torch.Generator().manual_seed(SEED + 8)
    f = cn01(gen, (N, N), dev).to(torch.complex128)
    x1 = torch.from_numpy(x1n.astype(np.float64)).to(dev)
    x2 = torch.from_numpy(x2n.astype(np.float64)).to(dev)
    y = torch.empty(M, dtype=torch.complex128, device=dev)
    plan = m.Plan(N, N, torch.complex128, device=dev)

    def step():
        plan.setpts(x1, x2)
        plan.type2(f, out=y)

    g = capture(step)
    s = time_graph(g, max(args.steps // 4, 10), 3)
    sel = torch.from_numpy(np.random.default_rng(SEED + 8).choice(M, 4096, replace=False)).to(dev)
    ref = direct_type2_subset(f, x1[sel], x2[sel])
    rel = ((y[sel] - ref).norm() / ref.norm()).item()
    d = cn01(gen, (M,), dev).to(torch.complex128)
    F = torch.empty(N, N, dtype=torch.complex128, device=dev)

    def step1():
        plan.setpts(x1, x2)
        plan.type1(d, out=F)

    g1 = capture(step1)
    s1 = time_graph(g1, max(args.steps // 4, 10), 3)
    rng = np.random.default_rng(SEED + 9)
    selk = torch.from_numpy(rng.choice(N * N, 256, replace=False)).to(dev)
    ref1 = direct_type1_subset(d, x1, x2, selk // N - N // 2, selk % N - N // 2, N, N)
    rel1 = ((F.reshape(-1)[selk] - ref1).norm() / ref1.norm()).item()
    return {"config": "complex128 type-2 NFFT, 512x512 modes, M=1048576 polar points, K=6, c=2; step = setpts + type-2",
            "value": M / s / 1e6, "unit": "Mpts/s", "ms_per_step": s * 1e3, "rel_l2": rel,
            "type1": {"config": "complex128 type-1 (the double NFFT2_2D_GPU), same points; step = setpts + type-1",
                      "value": M / s1 / 1e6, "unit": "Mpts/s", "ms_per_step": s1 * 1e3,
                      "rel_l2_256_modes": rel1}}


def bench_knab(args, dev):
    """knabNFFT2D (BP_FBP_FFBP/knabNFFT2D.m, the band-limited polar interpolation above NFFT1_2D in
    FBP.m:95): 512 x 512 (cos alpha x rho) samples of a band-limited test signal at half its Nyquist
    rate, interpolated at 1024 x 1024 points inside the input grid; eager calls (the wrapper reads
    the grid steps and extents on the host).  Accuracy: rel-L2 against the signal itself."""
    import miscellaneous_amd as m

    n, T1, T2 = 512, 0.05, 0.002
    f64 = torch.float64
    rho = 100 + T1 * torch.arange(n, dtype=f64, device=dev)
    cs = -0.1 + T2 * torch.arange(n, dtype=f64, device=dev)
    RHO, COS = torch.meshgrid(rho, cs, indexing="xy")
    B1, B2 = 0.5 / T1, 0.5 / T2
    rng = np.random.default_rng(SEED + 10)
    u = torch.from_numpy((rng.random(6) - 0.5) * B1 * 0.9).to(dev)
    v = torch.from_numpy((rng.random(6) - 0.5) * B2 * 0.9).to(dev)
    c = torch.from_numpy(rng.standard_normal(6) + 1j * rng.standard_normal(6)).to(dev)

    def sig(R, C):
        ph = 2 * np.pi * (R[..., None] * u + C[..., None] * v)
        return (torch.polar(torch.ones_like(ph), ph) * c).sum(-1)

    data = sig(RHO, COS).to(torch.complex64)
    RO, CO = torch.meshgrid(torch.linspace(float(rho[20]), float(rho[-21]), 1024, dtype=f64, device=dev),
                            torch.linspace(float(cs[20]), float(cs[-21]), 1024, dtype=f64, device=dev),
                            indexing="xy")

    def call(ro=RO, co=CO):
        return m.knabNFFT2D(0.0, RHO, COS, data, ro, co, B1, B2)

    def host_timed(fn, reps):
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return out, times

    for _ in range(3):
        out = call()
    torch.cuda.synchronize()
    reps = max(args.steps // 5, 20)
    # the same output-grid tensors every call: the grid statistics, points and tile sort are reused
    out, times = host_timed(call, reps)
    s = float(np.median(times))
    # a new output-grid tensor pair every call (copies made outside the timed region): statistics
    # (one host synchronisation), points and tile sort each time
    fresh = [(RO.clone(), CO.clone()) for _ in range(reps)]
    it = iter(fresh)
    _, times_new = host_timed(lambda: call(*next(it)), reps)
    del fresh
    # the type-2 inside, alone (1024^2 modes at the same 1M points), for the call's overhead
    from miscellaneous_amd import knab as mk
    g = mk._geoms[-1]
    plan = m.Plan(1024, 1024, torch.complex64, device=dev).setpts(g["x1"], g["x2"])
    F = cn01(torch.Generator().manual_seed(SEED + 11), (1024, 1024), dev)
    t2 = stage_ms(lambda: plan.type2(F), 20)
    t2s = stage_ms(lambda: plan.setpts_type2(g["x1"], g["x2"], F), 20)
    plan.close()
    ref = sig(RO, CO)
    rel = ((out.to(torch.complex128) - ref).norm() / ref.norm()).item()
    return {"config": "knabNFFT2D: 512x512 polar samples (B = 0.5/T in both variables, P = 18), 1024x1024 "
                      "output points; spectrum FFT + window + zero padding (a = b = 2) + type-2 NFFT on "
                      "1024x1024 modes; eager calls, host-timed one by one, on the same output-grid tensors "
                      "(geometry reused)",
            "value": RO.numel() / s / 1e6, "unit": "Mpts/s", "ms_per_call": s * 1e3,
            "ms_per_call_mean": float(np.mean(times)) * 1e3, "calls": reps,
            "ms_per_call_new_geometry": float(np.median(times_new)) * 1e3,
            "inner_type2_ms": t2, "inner_setpts_type2_ms": t2s,
            "rel_l2_vs_bandlimited_signal": rel}


def bench_pso(args, dev):
    """One PSO generation's far fields (Synthesis/PSO): 250 particles x NFFT2_2D_GPU on 88 x 88 with
    M = 1936 elements + the far-field filter, as one batched graph, against 250 drop-in calls."""
    import miscellaneous_amd as m

    P, Nu, Nv, M = 250, 44, 44, 1936
    rng = np.random.default_rng(SEED + 7)
    gen = m.FarFieldGeneration(P, Nu, Nv, M, device=dev)
    g = np.stack(np.meshgrid(np.arange(44), np.arange(44)), -1).reshape(-1, 2).astype(np.float64)
    xy = (g[None] - 21.5) * 1.9 + rng.uniform(-0.3, 0.3, (P, M, 2))
    gen.X.copy_(torch.from_numpy(xy[..., 0].astype(np.float32)))
    gen.Y.copy_(torch.from_numpy(xy[..., 1].astype(np.float32)))
    gen.excitations.copy_(cn01(torch.Generator().manual_seed(SEED + 7), (P, M), dev))
    gen.filter.copy_(torch.from_numpy(rng.uniform(0.5, 1.5, 4 * Nu * Nv).astype(np.float32)))
    graph = gen.capture()
    s = time_graph(graph, max(args.steps // 4, 10), 3)
    # the reference's way: one NFFT2_2D_GPU call (plus the filter) per particle
    out = torch.empty(2 * Nv, 2 * Nu, dtype=torch.complex64, device=dev)

    def loop():
        for p in range(P):
            m.NFFT2_2D_GPU(out, gen.excitations[p], gen.X[p], gen.Y[p], 2 * Nv, 2 * Nu, M)
            out.mul_(gen.filter.view(2 * Nv, 2 * Nu))

    loop()
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        loop()
    torch.cuda.synchronize()
    s_loop = (time.perf_counter() - t0) / reps
    return {"config": "far fields of one PSO generation (Synthesis/PSO): 250 particles, NFFT2_2D_GPU on 88x88 "
                      "(Nu = Nv = 44) from M = 1936 elements each, times the far-field filter; batched setpts + "
                      "type-1 + filter replayed from one graph",
            "ms_per_generation": s * 1e3, "value": P * M / s / 1e6, "unit": "Mpts/s",
            "dropin_loop_ms_per_generation": s_loop * 1e3, "speedup_vs_dropin_loop": s_loop / s}


def bench_1d(args, dev):
    """Batched 1-D NFFTs: the per-pulse type-2 of BP_FBP_FFBP/BP_NFFT.m:18 (64 pulses x 424 frequency
    samples at 65,536 pixel ranges each) and a type-1 batch (1024 lines of 4096 samples to 256 modes)."""
    import miscellaneous_amd as m

    gen = torch.Generator().manual_seed(SEED + 8)
    B, N, M = 64, 424, 65536
    f = cn01(gen, (B, N), dev)
    x = ((torch.rand(B, M, generator=gen) - 0.5) * 0.9 * N).to(dev)
    plan = m.Plan1D(N, device=dev)
    out = plan.type2(f, x)
    sel = torch.from_numpy(np.random.default_rng(SEED + 8).choice(M, 1024, replace=False)).to(dev)
    k = torch.arange(-N // 2, N // 2, device=dev, dtype=torch.float64)
    ph = torch.outer(x[0][sel].double(), k) / N
    ref = (torch.polar(torch.ones_like(ph), -2 * np.pi * (ph - torch.floor(ph))) @ f[0].to(torch.complex128))
    rel2 = ((out[0][sel].to(torch.complex128) - ref).norm() / ref.norm()).item()
    g2 = capture(lambda: plan.type2(f, x))
    s2 = time_graph(g2, max(args.steps // 4, 10), 3)
    B1, N1, M1 = 1024, 256, 4096
    d = cn01(gen, (B1, M1), dev)
    x1 = ((torch.rand(B1, M1, generator=gen) - 0.5) * N1).to(dev)
    plan1 = m.Plan1D(N1, device=dev)
    plan1.type1(d, x1)
    g1 = capture(lambda: plan1.type1(d, x1))
    s1 = time_graph(g1, max(args.steps // 4, 10), 3)
    return {"type2": {"config": f"{B} lines x {N} modes -> {M} points each (BP_NFFT.m per-pulse NFFT1_1D), complex64",
                      "ms": s2 * 1e3, "value": B * M / s2 / 1e6, "unit": "Mpts/s", "rel_l2_1024": rel2},
            "type1": {"config": f"{B1} lines x {M1} samples -> {N1} modes (NFFT2_1D), complex64, exact fixed-point sums",
                      "ms": s1 * 1e3, "value": B1 * M1 / s1 / 1e6, "unit": "Mpts/s"}}


def type3_workload(dev, n=1 << 20):
    """Type-3 inputs: n points uniform in [-50, 50]^2 with CN(0,1) weights, n outputs uniform in [-10, 10]^2
    (X1 S = Y1 T = 500: a 1636 x 1636 grid at eps 1e-6)."""
    rng = np.random.default_rng(SEED + 11)
    x = torch.from_numpy(rng.uniform(-50, 50, n)).to(dev)
    y = torch.from_numpy(rng.uniform(-50, 50, n)).to(dev)
    s = torch.from_numpy(rng.uniform(-10, 10, n)).to(dev)
    t = torch.from_numpy(rng.uniform(-10, 10, n)).to(dev)
    f = torch.from_numpy((rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)).to(dev)
    return x, y, s, t, f


def bench_type3(args, dev):
    """NUFFT_2D_GPU (the type-3 of NFFT/NFFT3/C/NFFT3_2D_C.cpp:199, FP64) at 1M inputs -> 1M outputs, eps 1e-6;
    host-timed whole calls (each includes the extent read-back that sizes the grid), rel-L2 on 128 outputs
    against an FP64 direct sum over all inputs."""
    import miscellaneous_amd as m

    x, y, s, t, f = type3_workload(dev)
    n = x.numel()
    eps = 1e-6
    par = m.nufft3_params(x, y, s, t, eps)
    F = m.NUFFT_2D(x, y, s, t, f, eps)
    torch.cuda.synchronize()
    reps = max(args.steps // 20, 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        m.NUFFT_2D(x, y, s, t, f, eps)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    sel = torch.from_numpy(np.random.default_rng(SEED + 12).choice(n, 128, replace=False)).to(dev)
    ref = torch.empty(128, dtype=torch.complex128, device=dev)
    for c in range(0, 128, 8):
        ph = torch.outer(s[sel[c:c + 8]], x) + torch.outer(t[sel[c:c + 8]], y)
        ref[c:c + 8] = torch.polar(torch.ones_like(ph), -ph) @ f
    rel = ((F[sel] - ref).norm() / ref.norm()).item()
    # the host-array form (NUFFT_2D_CPU's own contract): PCIe copies in and out included
    hx, hy, hs, ht, hf = (a.cpu().numpy() for a in (x, y, s, t, f))
    m.NUFFT_2D_host(hx, hy, hs, ht, hf, eps)
    t0 = time.perf_counter()
    for _ in range(3):
        m.NUFFT_2D_host(hx, hy, hs, ht, hf, eps)
    sec_host = (time.perf_counter() - t0) / 3
    return {"config": f"type-3 NUFFT (NUFFT_2D_CPU's algorithm, FP64), {n} inputs uniform in [-50, 50]^2 -> {n} "
                      f"outputs uniform in [-10, 10]^2, eps {eps}: grid {par['Mrx']} x {par['Mry']}, "
                      f"{2 * par['msp'] + 1}^2 Gaussian taps per point",
            "ms_per_call": sec * 1e3, "value": 2 * n / sec / 1e6, "unit": "Mpts/s (inputs + outputs)",
            "rel_l2_128": rel, "eps": eps,
            "host_arrays": {"ms_per_call": sec_host * 1e3, "value": 2 * n / sec_host / 1e6,
                            "note": "NUFFT_2D_GPU_host: host arrays in and out, the PCIe copies included"}}


def bench_non_pow2(args, dev):
    """Sizes that are not powers of two (chirp-z FFT passes): type-2 and type-1 at 1000 x 1000 modes against
    1024 x 1024, 1M uniform points, complex64, eager calls (points set once)."""
    import miscellaneous_amd as m

    res = {}
    gen = torch.Generator().manual_seed(SEED + 13)
    M = 1 << 20
    for N in (1000, 1024):
        x = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
        y = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
        f = cn01(gen, (N, N), dev)
        d = cn01(gen, (M,), dev)
        p = m.Plan(N, N, torch.complex64, device=dev).setpts(x, y)
        out = {}
        for name, fn in (("type2_ms", lambda: p.type2(f)), ("type1_ms", lambda: p.type1(d))):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(20):
                fn()
            torch.cuda.synchronize()
            out[name] = (time.perf_counter() - t0) / 20 * 1e3
        res[str(N)] = out
    res["config"] = "type-2 / type-1 NFFT, N x N modes, M = 1048576 uniform points, complex64; 1000 runs chirp-z passes"
    return res


def bench_ndft_dropin(args, dev):
    """NDFT2_2D_GPU on meshgrid outputs (the reference's caller passes the integer meshgrid): the FP64
    matrix-core contraction, at the PSO size (88 x 88, M = 1936) and at 256 x 256 with M = 65,536."""
    import miscellaneous_amd as m

    res = {}
    rng = np.random.default_rng(SEED + 9)
    for tag, (nu, M) in {"pso_88": (88, 1936), "grid_256": (256, 65536)}.items():
        X = torch.from_numpy(rng.uniform(-nu / 2, nu / 2, M)).to(dev)
        Y = torch.from_numpy(rng.uniform(-nu / 2, nu / 2, M)).to(dev)
        d = torch.from_numpy((rng.standard_normal(M) + 1j * rng.standard_normal(M)) / np.sqrt(2)).to(dev)
        n = torch.arange(nu * nu, device=dev)
        u = (n // nu - nu // 2).double()
        v = (n % nu - nu // 2).double()
        out = torch.empty(nu * nu, dtype=torch.complex128, device=dev)

        def call():
            m.NDFT2_2D_GPU(0, X, Y, u, v, d, out, nu, nu, M, nu * nu)

        call()
        torch.cuda.synchronize()
        reps = 20 if M < 10000 else 5
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        torch.cuda.synchronize()
        s = (time.perf_counter() - t0) / reps
        flops = 8.0 * M * nu * nu
        res[tag] = {"config": f"{nu}x{nu} meshgrid outputs, M = {M} points, FP64", "ms": s * 1e3,
                    "fp64_tflops": flops / s / 1e12, "mfma_util": flops / s / 78.6e12}
    res["note"] = "useful FP64 flops 8 M N over host-timed calls; peak 78.6 TFLOP/s FP64 matrix (spec)"
    return res


def bench_autograd(args, dev):
    """The torch.autograd.Function path of the gradient (forward, loss, backward to f and x), eager
    steps at 256 x 256, M = 262,144, next to the fused type2_lsq + lsq_adjoint step of bench_grad."""
    import miscellaneous_amd as m

    N, M = 256, 262144
    gen = torch.Generator().manual_seed(SEED + 10)
    x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev).requires_grad_(True)
    x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev).requires_grad_(True)
    f = cn01(gen, (N, N), dev).requires_grad_(True)
    y_obs = cn01(gen, (M,), dev)
    plan = m.Plan(N, N, torch.complex64, device=dev)

    def step():
        f.grad = x1.grad = x2.grad = None
        loss = ((m.nfft2d_type2(f, x1, x2, plan) - y_obs).abs() ** 2).sum()
        loss.backward()

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    reps = max(args.steps // 4, 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    s = (time.perf_counter() - t0) / reps

    # the same step (forward, loss, backward) captured whole into one graph and replayed
    def graphed_step():
        loss = ((m.nfft2d_type2(f, x1, x2, plan) - y_obs).abs() ** 2).sum()
        loss.backward()

    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        for _ in range(3):
            f.grad = x1.grad = x2.grad = None
            graphed_step()
    torch.cuda.current_stream(dev).wait_stream(side)
    torch.cuda.synchronize()
    f.grad = x1.grad = x2.grad = None
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        graphed_step()
    s_graph = time_graph(g, reps, 3)
    del g
    return {"config": "256x256 modes, M=262144 uniform; loss ||nfft2d_type2(f, x1, x2) - y||^2 through "
                      "torch.autograd (grad f and x); eager steps, the tile sort skipped for unchanged points",
            "steps_per_s": 1.0 / s, "ms_per_step": s * 1e3,
            "graphed": {"steps_per_s": 1.0 / s_graph, "ms_per_step": s_graph * 1e3,
                        "note": "forward, loss and backward captured into one graph (torch.cuda.graph), replayed"}}


def cpu_baseline(t2, seconds: float):
    """The CPU restatement of the reference algorithm on the host cores, median of 5 timed runs after a
    warm-up: FP32 type-2 on all cores (the reported value), and beside it the 1-thread run, the type-1
    (NFFT/NFFT2_2D.cu's direction, OpenMP atomic spread as NFFT/NFFT3/C/NFFT3_2D_C.cpp:78-81) and FP64."""
    import oracle

    f = t2["f"].cpu().numpy()
    x1, x2 = t2["x1n"], t2["x2n"]
    M, N = len(x1), f.shape[0]
    rng = np.random.default_rng(SEED)
    d = ((rng.standard_normal(M) + 1j * rng.standard_normal(M)) / np.sqrt(2)).astype(np.complex64)
    cores = oracle.num_threads()

    def median_time(fn, runs=5):
        fn()
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    legs = {}
    t = median_time(lambda: oracle.nfft1_2d(f, x1, x2, precision="float"))
    value = M / t / 1e6
    legs["type2_fp32"] = {"Mpts_s": value, "s": t, "threads": cores}
    legs["type1_fp32"] = {"Mpts_s": M / (s := median_time(lambda: oracle.nfft2_2d(d, x1, x2, N, N, precision="float"))) / 1e6,
                          "s": s, "threads": cores}
    legs["type2_fp64"] = {"Mpts_s": M / (s := median_time(lambda: oracle.nfft1_2d(f.astype(np.complex128), x1, x2))) / 1e6,
                          "s": s, "threads": cores}
    legs["type1_fp64"] = {"Mpts_s": M / (s := median_time(lambda: oracle.nfft2_2d(d.astype(np.complex128), x1, x2, N, N))) / 1e6,
                          "s": s, "threads": cores}
    oracle.set_num_threads(1)
    try:
        sub = slice(0, M // 8)                       # one eighth of the points: same grid work, 1/8 of the gathers
        s = median_time(lambda: oracle.nfft1_2d(f, x1[sub], x2[sub], precision="float"), runs=3)
        legs["type2_fp32_1thread_M_over_8"] = {"Mpts_s": (M // 8) / s / 1e6, "s": s, "threads": 1}
    finally:
        oracle.set_num_threads(cores)
    return dict(value=value, unit="Mpts/s", cores=cores, kind="port",
                sample=f"median of 5 full type-2 transforms (512x512, M=1048576 polar, FP32) of the CPU restatement "
                       f"of BP_FBP_FFBP/NFFT1_2D.m on {cores} threads, after one warm-up; legs: type-1 "
                       f"(NFFT/NFFT2_2D.cu), FP64, and 1 thread on M/8 points", legs=legs)


def measure_copy_bandwidth(dev, nbytes=1 << 30, reps=10):
    """A plain device-to-device copy of 1 GiB, timed with HIP events: (read + write bytes) / time, in GB/s,
    the STREAM-like figure beside the 8 TB/s spec."""
    src = torch.empty(nbytes, dtype=torch.uint8, device=dev).fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    return gbs


def profile_workload(args):
    """A few eager calls of one workload, for rocprofv3 counter passes (no timing of its own)."""
    import miscellaneous_amd as m

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    gen = torch.Generator().manual_seed(SEED)
    if args.profile == "copy":
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
    elif args.profile == "type3":
        x, y, s, t, f = type3_workload(dev)
        for _ in range(args.steps):
            m.NUFFT_2D(x, y, s, t, f, 1e-6)
    elif args.profile == "pso":
        P, Nu, Nv, M = 250, 44, 44, 1936
        rng = np.random.default_rng(SEED + 7)
        gen = m.FarFieldGeneration(P, Nu, Nv, M, device=dev)
        g = np.stack(np.meshgrid(np.arange(44), np.arange(44)), -1).reshape(-1, 2).astype(np.float64)
        xy = (g[None] - 21.5) * 1.9 + rng.uniform(-0.3, 0.3, (P, M, 2))
        gen.X.copy_(torch.from_numpy(xy[..., 0].astype(np.float32)))
        gen.Y.copy_(torch.from_numpy(xy[..., 1].astype(np.float32)))
        gen.excitations.copy_(cn01(torch.Generator().manual_seed(SEED + 7), (P, M), dev))
        gen.filter.copy_(torch.from_numpy(rng.uniform(0.5, 1.5, 4 * Nu * Nv).astype(np.float32)))
        graph = gen.capture()
        for _ in range(args.steps):
            graph.replay()
    elif args.profile == "batch":
        sizes, rng = batch_sizes(SEED + 3)
        fs = [cn01(gen, (n1, n2), dev) for (n1, n2, _) in sizes]
        pts = [batch_points(rng, n1, mm) for (n1, _, mm) in sizes]
        outs = [torch.empty(mm, dtype=torch.complex64, device=dev) for (_, _, mm) in sizes]
        b = m.Batch(sizes, device=dev).set_points([torch.from_numpy(a).to(dev) for a, _ in pts],
                                                  [torch.from_numpy(c).to(dev) for _, c in pts]).set_io(fs, outs)
        for _ in range(args.steps):
            b.setpts_type2()
    elif args.profile == "ndft_dropin":
        nu, M = 256, 65536
        rng = np.random.default_rng(SEED + 9)
        X = torch.from_numpy(rng.uniform(-nu / 2, nu / 2, M)).to(dev)
        Y = torch.from_numpy(rng.uniform(-nu / 2, nu / 2, M)).to(dev)
        d = torch.from_numpy((rng.standard_normal(M) + 1j * rng.standard_normal(M)) / np.sqrt(2)).to(dev)
        n = torch.arange(nu * nu, device=dev)
        u, v = (n // nu - nu // 2).double(), (n % nu - nu // 2).double()
        out = torch.empty(nu * nu, dtype=torch.complex128, device=dev)
        for _ in range(args.steps):
            m.NDFT2_2D_GPU(0, X, Y, u, v, d, out, nu, nu, M, nu * nu)
    elif args.profile == "ndft_bf16x3":
        N, M = 256, 65536
        f = cn01(gen, (N, N), dev)
        x1 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
        x2 = ((torch.rand(M, generator=gen) - 0.5) * N).to(dev)
        p = m.NdftPlan(N, N, "bf16x3", device=dev).set_modes(f)
        out = torch.empty(M, dtype=torch.complex64, device=dev)
        for _ in range(args.steps):
            p.eval(x1, x2, out=out)
    else:
        N = 512
        x1n, x2n = polar_points(N)
        plan = m.Plan(N, N, torch.complex64, device=dev).setpts(torch.from_numpy(x1n).to(dev),
                                                                 torch.from_numpy(x2n).to(dev))
        f = cn01(gen, (N, N), dev)
        d = cn01(gen, (len(x1n),), dev)
        for _ in range(args.steps):
            plan.type2(f) if args.profile == "type2" else plan.type1(d)
    torch.cuda.synchronize()
    print(f"profile workload {args.profile}: {args.steps} calls done", flush=True)


# The workloads of the JSON line beside the headline and the gradient, in run order.
EXTRA_LEGS = {
    "batch": bench_batch,
    "ndft": bench_ndft,
    "large_1024": bench_large,
    "type1_polar512": bench_type1,
    "fp64_polar512": bench_fp64,
    "knab_polar_interp": bench_knab,
    "pso_generation": bench_pso,
    "nfft_1d": bench_1d,
    "ndft2_2d_dropin": bench_ndft_dropin,
    "autograd": bench_autograd,
    "type3": bench_type3,
    "non_pow2": bench_non_pow2,
}


def main():
    faulthandler.enable()          # a crash prints the Python stack to stderr
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batch, NDFT and 1024^2 workloads (headline and gradient only)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "gather_polar512_traffic.json"),
                    help="JSON with rocprofv3-measured HBM bytes per gather launch")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (to scale FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--profile", choices=["type2", "type1", "ndft_bf16x3", "ndft_dropin", "pso", "copy", "type3",
                                          "batch"],
                    help="run only --steps eager calls of one workload (polar 512^2 type-2 / type-1, the "
                         "BF16x3 NDFT at 256^2, M = 65536, the FP64 type-3 at 1M -> 1M, or a 256 MiB device "
                         "copy) and exit: the command "
                         "the rocprofv3 counter runs profile")
    ap.add_argument("--only", choices=sorted(EXTRA_LEGS),
                    help="run one of the extra workloads alone and print its JSON object")
    args = ap.parse_args()
    if args.profile:
        profile_workload(args)
        return
    if args.only:
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        print(json.dumps({args.only: EXTRA_LEGS[args.only](args, dev)}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t2 = bench_type2(args, dev)
    gr = None if args.no_grad else bench_grad(args, dev)
    extra = {}
    if not args.no_extra:
        for name, leg in EXTRA_LEGS.items():
            extra[name] = leg(args, dev)
    cpu = None if args.no_cpu_baseline else cpu_baseline(t2, args.cpu_seconds)

    M = t2["M"]
    achieved = t2["gather_bytes"] / (t2["gather_ms"] * 1e-3) / 1e9
    traffic, tj = None, {}
    if os.path.exists(args.traffic):
        with open(args.traffic) as fh:
            tj = json.load(fh)
        traffic = tj.get("bytes_per_launch")
    secondary = None
    if tj.get("valu_insts_per_launch"):
        # VALU issue floor: a wave-64 FP32 VALU instruction occupies its SIMD 4 cycles; 1024 SIMDs at the
        # 2.4 GHz peak engine clock (no packed FP32 here, so 16 lanes per SIMD per cycle)
        issue_us = tj["valu_insts_per_launch"] * 4 / (1024 * 2.4e3)
        secondary = {"bound": "valu_issue", "valu_insts_per_launch": tj["valu_insts_per_launch"],
                     "floor_us": issue_us, "achieved_us": t2["gather_ms"] * 1e3,
                     "frac": issue_us / (t2["gather_ms"] * 1e3),
                     "note": "rocprofv3 SQ_INSTS_VALU of the same kernel and workload x 4 cycles / (1024 SIMDs x 2.4 GHz)"}
    copy_gbs = measure_copy_bandwidth(dev)
    out = {
        "metric": "2D type-2 NFFT Mpts/s at rel-L2<=1e-5 (512^2, M=1M polar)",
        "value": M / t2["s_full"] / 1e6,
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t2["s_full"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded CN(0,1) modes, polar point grid)",
        "config": {"workload": "type-2 NFFT, 512x512 modes, M=1048576 polar points (512 angles x 2048 radii), "
                               "K=6, c=2, complex64; step = setpts + type-2 (tile sort beside the grid passes)",
                   "N1": 512, "N2": 512, "M": M, "K": 6, "c": 2},
        "rel_l2": t2["rel_l2"],
        "exec_only": {"value": M / t2["s_exec"] / 1e6, "unit": "Mpts/s", "ms_per_step": t2["s_exec"] * 1e3},
        "one_stream": {"value": M / t2["s_serial"] / 1e6, "unit": "Mpts/s", "ms_per_step": t2["s_serial"] * 1e3,
                       "note": "setpts and type-2 serialised on one stream (the step overlaps the tile sort "
                               "with the grid passes)"},
        "stage_ms": {"setpts": t2["setpts_ms"], "grid_a_b": t2["grid_ms"], "gather_c": t2["gather_ms"]},
        "stage_roofline": {"a_b_precomp_pad_fft": stage_roofline(168 * t2["N"] * t2["N"], t2["grid_ms"]),
                           "c_gather": stage_roofline(t2["gather_bytes"], t2["gather_ms"])},
        "type2_hbm_frac": t2["type2_bytes"] / t2["s_exec"] / (HBM_PEAK_GBS * 1e9),
        "roofline": {"kernel": "k_gather_pairs<2> (stage c)", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": t2["gather_bytes"], "measured_copy_GBs": copy_gbs,
                     "secondary": secondary},
    }
    if gr is not None:
        out["grad"] = {"steps_per_s": 1.0 / gr["s"], "ms_per_step": gr["s"] * 1e3,
                       "config": "256x256 modes, M=262144 uniform, loss ||A(x)f - y||^2, grad f and x, GD update; "
                                 "step = type2_lsq (forward, residual and grad x in one gather; the residual stays in the plan) + lsq_adjoint (its adjoint, spread in tile order) + update",
                       "hbm_frac": gr["bytes"] / gr["s"] / (HBM_PEAK_GBS * 1e9),
                       "loss_first": gr["loss0"], "loss_after_steps": gr["loss_final"], "gd_steps": gr["steps"],
                       "gd_lr": gr["lr"]}
    out.update(extra)
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
