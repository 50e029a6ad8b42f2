#!/usr/bin/env python3
"""Benchmark of the device subtask path on one MI355X.

Metric: tasks/s (+ p50/p99 task latency) of the served + batched HIP stub, next to the
per-subtask-launch paths on the same GPU and the reference CPU algorithm on the host cores.

Workload (default `imagefilter`): imageFiltering of one 8192 x 8192 RGB uint8 image, 5 x 5
integer weights in [-4, 4], sum >> 4 clamped to [0, 255], split into 16 subtasks of 2048 x 2048
tiles with 2-pixel halos. Inputs are synthetic (seeded) and resident in HBM before the timed
region. One step = one task (submit -> all 16 subtasks done).

Roofline: the dominant kernel is the persistent filter kernel of the batched flush; its
algorithmic traffic per launch is 2 * W * H * 3 bytes (each input and output byte once), divided
by the flush time measured with HIP events on the stream the kernel runs on.

A second section serves a seeded Poisson mix of matrixMultiply / imageFiltering / pageRank
requests (BATCHED vs PLAIN stub on the same stream) and reports tasks/s and p50/p99 latency.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                       [--serving-requests N] [--serving-rate R] [--serving-slots S]
Defaults of K / W per workload: imagefilter and matmul 200 / 3 (a task takes 0.1-0.2 ms, so 200
keep one slow task from moving the mean), pagerank 10 / 1, lu 5 / 1.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# The dominant kernel's duration comes from HIP events on its stream around every KT_EVERY-th
# flush of the timed region: each event pair costs the device a few microseconds, which timing
# every flush would add to the tasks being measured.
KT_EVERY = 4


def percentile(v, q):
    v = np.asarray(v, dtype=np.float64)
    return float(np.percentile(v, q)) if v.size else None


def time_tasks(run, steps):
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        lat.append(run())
    t1 = time.perf_counter()
    return t1 - t0, lat


# Untimed tasks run before the warm-up steps, in seconds: under sustained load the GPU's clocks
# settle within a few milliseconds (measured on the imageFiltering task: its kernel takes about
# 0.167 ms over the first tens of tasks and 0.150 ms after), so the timed steps see the
# steady state that a serving loop runs in.
PREWARM_S = 0.3


def prewarm(run, seconds=PREWARM_S):
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        run()
        n += 1
    return n


def traffic_from_profiles(name):
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("traffic_bytes_per_launch")
    except Exception:
        return None


def bench_imagefilter(args, U):
    H = W = 8192
    r, shift, tile = 2, 4, 2048
    rng = np.random.default_rng(20240601)
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    weights = rng.integers(-4, 5, (2 * r + 1) ** 2).astype(np.int8)
    f = U.ImageFilter(W, H, weights, r, mode=1, shift=shift, tile=tile)
    f.set_input(img)
    algo_bytes = 2 * W * H * 3

    # served + batched: HBM-resident address spaces, one persistent launch per task, hipGraph
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    U.set_kernel_timing(True, every=KT_EVERY)
    warm = prewarm(lambda: f.run(U.DEVICE_GPU))
    for _ in range(args.warmup):
        f.run(U.DEVICE_GPU)
    U.device_sync()
    U.kernel_time(reset=True)
    U.reset_stats()
    elapsed, lat = time_tasks(lambda: f.run(U.DEVICE_GPU), args.steps)
    U.device_sync()
    kms, flushes = U.kernel_time(reset=True)
    st = U.stats()
    U.set_kernel_timing(False)
    batched = {"tasks_per_s": args.steps / elapsed, "p50_ms": 1e3 * percentile(lat, 50),
               "p99_ms": 1e3 * percentile(lat, 99)}
    kernel_ms = kms / max(1, flushes)

    # per-subtask-launch comparators on the same GPU
    comps = {}
    for name, mode in (("plain", U.PM_HIP_MODE_PLAIN), ("resident", U.PM_HIP_MODE_RESIDENT)):
        U.set_hip_mode(mode)
        f.run(U.DEVICE_GPU)
        n = max(2, min(args.steps, args.comparator_steps))
        e, l = time_tasks(lambda: f.run(U.DEVICE_GPU), n)
        comps[name] = {"tasks_per_s": n / e, "p50_ms": 1e3 * percentile(l, 50), "p99_ms": 1e3 * percentile(l, 99)}
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED)
    f.close()

    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("imagefilter_8192"),
                "kernel": "k_imagefilter_persistent", "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes,
                "algorithmic_bytes_per_launch": algo_bytes}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_imagefilter(img, weights, r, shift, tile)

    return {
        "metric": "tasks/s (served+batched, imageFiltering 8192^2 task)",
        "value": round(batched["tasks_per_s"], 3),
        "unit": "tasks/s",
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "dtype": "u8",
        "config": {"workload": "imageFiltering 8192x8192 RGB u8, 5x5 int weights in [-4,4], >>4 + clamp, "
                               "16 subtasks of 2048x2048 tiles with 2-px halo",
                   "hip_mode": "batched+graph", "tile": tile, "radius": r},
        "p50_ms": round(batched["p50_ms"], 3), "p99_ms": round(batched["p99_ms"], 3),
        "per_subtask_launch": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in comps.items()},
        "speedup_vs_plain": round(batched["tasks_per_s"] / comps["plain"]["tasks_per_s"], 2),
        "roofline": roofline,
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
        "stub_stats": {k: st[k] for k in ("subtasks", "batches", "h2d_bytes", "d2h_bytes", "persistent_launches")},
    }


def bench_matmul(args, U):
    """matrixMultiply N=2048 float32, A and B uniform in [-1, 1) from the seed, 16 subtasks of
    512 x 512 output blocks; split-BF16 MFMA products (3 per fp32 product) with fp32 accumulation.
    One step = one task. Roofline: MFMA-issued bf16 flop (3 * 2 N^3) per batched flush over the
    flush's device time, against the dense bf16 peak."""
    n, bs = 2048, 512
    rng = np.random.default_rng(20240601)
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    B = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    mm.set_inputs(A, B)
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    U.set_kernel_timing(True, every=KT_EVERY)
    warm = prewarm(lambda: mm.run(U.DEVICE_GPU))
    for _ in range(args.warmup):
        mm.run(U.DEVICE_GPU)
    U.device_sync()
    U.kernel_time(reset=True)
    elapsed, lat = time_tasks(lambda: mm.run(U.DEVICE_GPU), args.steps)
    U.device_sync()
    kms, flushes = U.kernel_time(reset=True)
    U.set_kernel_timing(False)
    kernel_ms = kms / max(1, flushes)
    C64 = A.astype(np.float64) @ B.astype(np.float64)
    err = float(np.linalg.norm(mm.output() - C64) / np.linalg.norm(C64))
    comps = {}
    for name, mode in (("plain", U.PM_HIP_MODE_PLAIN), ("resident", U.PM_HIP_MODE_RESIDENT)):
        U.set_hip_mode(mode)
        mm.run(U.DEVICE_GPU)
        k = max(2, min(args.steps, args.comparator_steps))
        e, l = time_tasks(lambda: mm.run(U.DEVICE_GPU), k)
        comps[name] = {"tasks_per_s": round(k / e, 3), "p50_ms": round(1e3 * percentile(l, 50), 3)}
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED)
    mm.close()
    flop = 2.0 * n ** 3
    issued = 3 * flop
    achieved = issued / (kernel_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle_lib as O
        cores = os.cpu_count() or 1
        k = min(16, cores)
        C = np.zeros((n, n), dtype=np.float32)
        t0 = time.perf_counter()
        O.matmul_blocks(A, B, bs, 0, k, threads=k, C_out=C)
        dt = time.perf_counter() - t0
        cpu = {"value": round((k / 16) / dt, 4), "unit": "tasks/s", "cores": k, "kind": "port",
               "sample": f"{k} of the 16 subtasks (plain i-j-k float loop, one thread each), {dt:.2f} s"}
    return {
        "metric": "tasks/s (served+batched, matrixMultiply N=2048 fp32 task)",
        "value": round(args.steps / elapsed, 3), "unit": "tasks/s",
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "dtype": "f32 (split-bf16 x3 MFMA)",
        "config": {"workload": "matrixMultiply N=2048 float32, 16 subtasks of 512x512", "hip_mode": "batched+graph"},
        "p50_ms": round(1e3 * percentile(lat, 50), 3), "p99_ms": round(1e3 * percentile(lat, 99), 3),
        "per_subtask_launch": comps, "fp64_normwise_rel_err": err,
        "fp32_equivalent_tflops": round(flop / (kernel_ms * 1e-3) / 1e12, 1),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 4), "traffic": traffic_from_profiles("matmul_2048"),
                     "kernel": "k_gemm_persistent",
                     "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes, "algorithmic_flop_per_launch": issued},
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
    }


def bench_pagerank(args, U):
    """pageRank on a 10M-page synthetic power-law web graph (out-degree P(k) ~ (k + k0)^-2 on
    [1, 100], mean 10, 5 % dangling, preferential targets), 20 iterations, 1M-page subtasks.
    One step = one pageRank computation (init task + 20 iteration tasks) over the source-blocked
    graph layout. Roofline: per iteration 12 N + 4 E bytes (row pointers, degree, rank in / out,
    in-link ids) over the flush's device time (the blocked gather and the combine kernel). The
    plain in-link CSR layout of the same graph is timed beside it."""
    pages, per, iters = 10_000_000, 1_000_000, 20
    rp, col = U.webgen(pages, 20240601)
    edges = int(col.size)
    pr = U.PageRank(rp, col, per)
    if not pr.blocked:
        raise RuntimeError("bench: the 10M-page graph did not take the source-blocked layout")
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    warm = prewarm(lambda: pr.run(iters, U.DEVICE_GPU))
    for _ in range(args.warmup):
        pr.run(iters, U.DEVICE_GPU)
    U.device_sync()
    U.set_kernel_timing(True, every=KT_EVERY)
    U.kernel_time(reset=True)
    elapsed, lat = time_tasks(lambda: pr.run(iters, U.DEVICE_GPU), args.steps)
    U.device_sync()
    kms, flushes = U.kernel_time(reset=True)
    U.set_kernel_timing(False)
    y = pr.ranks()
    kernel_ms = kms / max(1, flushes)
    comps = {}
    U.set_hip_mode(U.PM_HIP_MODE_PLAIN)
    e, l = time_tasks(lambda: pr.run(iters, U.DEVICE_GPU), 1)
    comps["plain"] = {"tasks_per_s": round(1 / e, 4), "p50_ms": round(1e3 * l[0], 1)}
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED)
    pr.close()
    # the same graph in the plain in-link CSR layout (CSR-stream pull, no source blocking)
    U.pr_set_layout(False)
    prc = U.PageRank(rp, col, per)
    U.pr_set_layout(True)
    prc.run(iters, U.DEVICE_GPU)
    U.device_sync()
    U.set_kernel_timing(True, every=KT_EVERY)
    U.kernel_time(reset=True)
    e, l = time_tasks(lambda: prc.run(iters, U.DEVICE_GPU), 1)
    U.device_sync()
    cms, cfl = U.kernel_time(reset=True)
    U.set_kernel_timing(False)
    comps["csr_layout"] = {"tasks_per_s": round(1 / e, 4), "kernel_ms_per_iteration": round(cms / max(1, cfl), 4)}
    prc.close()
    # the reference's decomposition: source-page subtasks push into private full-length vectors
    # that REDUCE_ADD folds in HBM (records padded to the largest out-degree, 101 uint32 per page)
    prp = U.PageRankPush(rp, col, per)
    prp.run(1, U.DEVICE_GPU)
    U.device_sync()
    e, l = time_tasks(lambda: prp.run(iters, U.DEVICE_GPU), 1)
    yp = prp.ranks()
    prp.close()
    comps["push_reduce_add"] = {"tasks_per_s": round(1 / e, 4),
                                "rel_l1_vs_pull": float(np.abs(yp.astype(np.float64) - y).sum() / np.abs(y).sum())}
    it_bytes = 12 * pages + 4 * edges
    achieved = it_bytes / (kernel_ms * 1e-3) / 1e9
    cpu, err = None, None
    if not args.no_cpu_baseline:
        from oracle import oracle_lib as O
        t0 = time.perf_counter()
        yo = O.pagerank(rp, col, iters, per)
        dt = time.perf_counter() - t0
        err = float(np.abs(y.astype(np.float64) - yo).sum() / np.abs(yo).sum())
        cpu = {"value": round(1 / dt, 4), "unit": "tasks/s", "cores": 1, "kind": "port",
               "sample": f"one full task (20 iterations, 10 push subtasks + REDUCE_ADD each), one thread, {dt:.1f} s"}
    return {
        "metric": "tasks/s (served+batched, pageRank 10M pages x 20 iterations)",
        "value": round(args.steps / elapsed, 3), "unit": "tasks/s",
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "dtype": "f32",
        "config": {"workload": f"pageRank {pages} pages, {edges} links, 20 iterations, 1M-page subtasks",
                   "hip_mode": "batched+graph"},
        "p50_ms": round(1e3 * percentile(lat, 50), 3), "p99_ms": round(1e3 * percentile(lat, 99), 3),
        "per_subtask_launch": comps, "rel_l1_vs_oracle": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("pagerank_10m"),
                     "kernel": "k_pagerank_blocked_persistent + k_pagerank_combine_persistent",
                     "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes,
                     "algorithmic_bytes_per_launch": it_bytes},
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
    }


def bench_lu(args, U):
    """luDecomposition N=16384 float32, A = U[0,1) + N I (column-major), 512 x 512 blocks, the
    trailing update in split-BF16 MFMA, then A x = b for 8 seeded right-hand sides with
    FP64-residual iterative refinement. One step = factorisation + refined solve, with the
    matrix resident in HBM. Errors are against the FP64 truth (b = A x_true in FP64); the plain
    FP32 path (VALU trailing update) is compared at N=4096."""
    n, bs, nrhs = 16384, 512, 8
    rng = np.random.default_rng(20240601)
    M = rng.random((n, n), dtype=np.float32)
    M[np.diag_indices(n)] += np.float32(n)
    A = M.T                                   # column-major view of a U[0,1) + N I matrix
    x_true = rng.uniform(-1, 1, (n, nrhs))
    b = np.asfortranarray(M.T.astype(np.float64) @ x_true)
    lu = U.LU(n, bs, 0)
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    fact, solve, infos = [], [], []
    for step in range(args.warmup + args.steps):
        lu.set_matrix(A)
        lu.prefetch()
        U.device_sync()
        t0 = time.perf_counter()
        lu.factor(U.DEVICE_GPU)
        t1 = time.perf_counter()
        x, info = lu.solve(b, 5, U.DEVICE_GPU)
        t2 = time.perf_counter()
        if step >= args.warmup:
            fact.append(t1 - t0)
            solve.append(t2 - t1)
            infos.append(info)
        print(f"lu: step {step + 1}/{args.warmup + args.steps} factor {t1 - t0:.4f} s solve {t2 - t1:.4f} s",
              file=sys.stderr, flush=True)
    fwd = float(np.abs(x - x_true).max() / np.abs(x_true).max())
    lu.close()
    del M, A
    factor_s, solve_s = float(np.median(fact)), float(np.median(solve))
    flop = 2.0 / 3.0 * n ** 3
    trailing_issued = 3 * 2.0 * sum((n - (k + 1) * bs) ** 2 * bs for k in range(n // bs - 1))
    # plain FP32 comparison at N=4096 (VALU trailing update) against split-BF16 at the same size
    n2 = 4096
    M2 = rng.random((n2, n2), dtype=np.float32)
    M2[np.diag_indices(n2)] += np.float32(n2)
    x2 = rng.uniform(-1, 1, (n2, nrhs))
    b2 = np.asfortranarray(M2.T.astype(np.float64) @ x2)
    small = {}
    for name, prec in (("split_bf16", 0), ("plain_fp32", 2)):
        print(f"lu: N={n2} comparison, {name}", file=sys.stderr, flush=True)
        l2 = U.LU(n2, bs, prec)
        l2.set_matrix(M2.T)
        l2.prefetch()
        t0 = time.perf_counter()
        l2.factor(U.DEVICE_GPU)
        t1 = time.perf_counter()
        F = l2.factors()
        y, inf2 = l2.solve(b2, 5, U.DEVICE_GPU)
        l2.close()
        L = np.tril(F.astype(np.float64))
        Uu = np.triu(F.astype(np.float64), 1) + np.eye(n2)
        A2 = M2.T.astype(np.float64)
        small[name] = {"factor_s": round(t1 - t0, 4),
                       "factor_backward_err": float(np.linalg.norm(A2 - L @ Uu) / np.linalg.norm(A2)),
                       "refine_steps": inf2["steps"], "solve_backward_err": inf2["backward_error"],
                       "solve_first_step_backward_err": inf2["backward_error_first"],
                       "forward_err": float(np.abs(y - x2).max() / np.abs(x2).max())}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle_lib as O
        n3 = 2048
        print(f"lu: CPU baseline at N={n3}", file=sys.stderr, flush=True)
        M3 = np.asfortranarray(rng.random((n3, n3), dtype=np.float32) + n3 * np.eye(n3, dtype=np.float32))
        t0 = time.perf_counter()
        O.lu(M3)
        dt = time.perf_counter() - t0
        est = dt * (n / n3) ** 3
        cpu = {"value": round(1.0 / est, 6), "unit": "tasks/s", "cores": 1, "kind": "port",
               "sample": f"unblocked LU (sscal + sger order) at N={n3} on one thread, {dt:.2f} s, "
                         f"scaled by (N/{n3})^3 to N={n} (factorisation only)"}
    return {
        "metric": "tasks/s (luDecomposition N=16384 factorisation + 8-RHS refined solve)",
        "value": round(1.0 / (factor_s + solve_s), 4), "unit": "tasks/s",
        "ms_per_step": round(1e3 * (factor_s + solve_s), 1), "dtype": "f32 (split-bf16 x3 MFMA update) + f64 residual",
        "config": {"workload": "luDecomposition N=16384 fp32 U[0,1)+N*I, 512 blocks, 8 RHS refinement",
                   "hip_mode": "batched+graph"},
        "factor_s": round(factor_s, 4), "solve_s": round(solve_s, 4),
        "factor_fp32_equivalent_tflops": round(flop / factor_s / 1e12, 2),
        "refine_steps": infos[-1]["steps"], "solve_backward_err": infos[-1]["backward_error"],
        "forward_err": fwd, "n4096_comparison": small,
        "roofline": {"bound": "mfma", "achieved": round(trailing_issued / factor_s / 1e12, 1), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(trailing_issued / factor_s / 1e12 / 2500.0, 4),
                     "traffic": None, "kernel": "whole factorisation (trailing-update MFMA flop / factor time)",
                     "algorithmic_flop_per_launch": trailing_issued},
        "cpu_baseline": cpu,
    }


def bench_serving(args, U):
    """Serving mix on one GPU: a seeded Poisson stream of matrixMultiply (N log-uniform in
    256..4096, float32, split-BF16 MFMA), imageFiltering (RGB u8 side uniform in 512..4096, 5x5,
    >>4 + clamp) and pageRank (100k..2M pages, 20 iterations) requests over resident datasets,
    each request's output returned to the host. The same stream is served by the BATCHED stub
    (continuous batching into persistent kernels, hipGraphs) and by the PLAIN per-subtask-launch
    stub; both report tasks/s and p50/p99 latency (arrival -> result on the host)."""
    S = U.Server()
    t0 = time.perf_counter()
    mm = [256, 512, 768, 1024, 1536, 2048, 3072, 4096]
    cat = {"mm": [(n, S.add_mm(n, 512 if n % 512 == 0 else 256, 0, seed=100 + i)) for i, n in enumerate(mm)],
           "if": [(s_, S.add_if(s_, 2, 1, 4, 2048, seed=200 + i)) for i, s_ in enumerate(range(512, 4097, 512))],
           "pr": [(p_, S.add_pr(p_, 1_000_000, 20, seed=300 + i))
                  for i, p_ in enumerate([100_000, 500_000, 1_000_000, 2_000_000])]}
    setup_s = time.perf_counter() - t0
    n, rate, slots = args.serving_requests, args.serving_rate, args.serving_slots
    ds, arr = U.serving_stream(20240601, n, rate, cat)
    kind_of = {}
    for k, entries in cat.items():
        for _, d in entries:
            kind_of[d] = k
    kinds = np.array([kind_of[int(d)] for d in ds])
    out = {"requests": n, "rate_per_s": rate, "slots": slots, "seed": 20240601,
           "mix": {k: int((kinds == k).sum()) for k in ("mm", "if", "pr")},
           "results_to_host": True, "dataset_setup_s": round(setup_s, 2)}
    warm = np.array(sorted({int(d) for d in ds}), dtype=np.int32)
    for name, mode in (("batched", U.PM_HIP_MODE_BATCHED), ("plain", U.PM_HIP_MODE_PLAIN)):
        U.set_hip_mode(mode, graphs=True)
        S.run(warm, np.zeros(len(warm)), 4, U.DEVICE_GPU)          # uploads, clones, graphs
        lat, _, wall = S.run(ds, arr, slots, U.DEVICE_GPU)
        res = {"tasks_per_s": round(n / wall, 2), "wall_s": round(wall, 3),
               "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3)}
        res["per_class"] = {k: {"p50_ms": round(percentile(lat[kinds == k], 50), 3),
                                "p99_ms": round(percentile(lat[kinds == k], 99), 3)}
                            for k in ("mm", "if", "pr") if (kinds == k).any()}
        out[name] = res
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    out["speedup_vs_plain"] = round(out["batched"]["tasks_per_s"] / out["plain"]["tasks_per_s"], 2)
    S.close()
    return out


def copy_calibration(nbytes, reps=3):
    """D2D copies of `nbytes` (the size of one image): the reference point for the HBM counters
    of a profiled run (FETCH_SIZE / WRITE_SIZE of a copy of known size)."""
    from unicorn_amd import hip as H
    L = H.lib()
    src, dst = H.DeviceBuffer(nbytes), H.DeviceBuffer(nbytes)
    for _ in range(reps):
        H.check(L.uh_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "uh_memcpy_d2d")
    H.check(L.uh_device_sync(), "sync")
    src.free()
    dst.free()


def cpu_baseline_imagefilter(img, weights, r, shift, tile):
    """The oracle (C restatement of the reference CPU callback) on the host cores: one full task
    (16 tile subtasks), one worker thread per core pulling tiles, as the reference's CPU stubs."""
    from oracle import oracle_lib as O
    cores = os.cpu_count() or 1
    H, W, _ = img.shape
    ntiles = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    out = np.zeros_like(img)
    t0 = time.perf_counter()
    O.imagefilter_tiled(img, weights, r, 1, shift, tile, 0, ntiles, cores, out=out)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "tasks/s", "cores": cores, "kind": "port",
            "sample": f"one full task ({ntiles} tiles of {tile}x{tile}) on {cores} threads, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="imagefilter", choices=["imagefilter", "matmul", "pagerank", "lu"])
    ap.add_argument("--comparator-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=300, help="0 skips the serving mix")
    ap.add_argument("--serving-rate", type=float, default=1000.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run D2D copies of one image size (reference for HBM counters)")
    args = ap.parse_args()
    steps, warmup = {"imagefilter": (200, 3), "matmul": (200, 3), "pagerank": (10, 1), "lu": (5, 1)}[args.workload]
    if args.steps is None:
        args.steps = steps
    if args.warmup is None:
        args.warmup = warmup
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}), flush=True)
        return 2
    import unicorn_amd as U
    U.init(0)
    if not U.has_gpu():
        print(json.dumps({"error": "no HIP device visible"}), flush=True)
        return 1
    if args.copy_calibration:
        copy_calibration(8192 * 8192 * 3)
    res = {"imagefilter": bench_imagefilter, "matmul": bench_matmul, "pagerank": bench_pagerank,
           "lu": bench_lu}[args.workload](args, U)
    line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic (seeded)", "config": res.pop("config")}
    line.update(res)
    if args.serving_requests > 0 and args.workload == "imagefilter":
        line["serving"] = bench_serving(args, U)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
