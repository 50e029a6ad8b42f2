#!/usr/bin/env python3
"""Benchmark of the device subtask path on one MI355X.

Metric: tasks/s (+ p50/p99 task latency) of the served + batched HIP stub, next to the
per-subtask-launch paths on the same GPU and the reference CPU algorithm on the host cores.

Headline workload: imageFiltering of one 8192 x 8192 RGB uint8 image, 5 x 5 integer weights in
[-4, 4], sum >> 4 clamped to [0, 255], split into 16 subtasks of 2048 x 2048 tiles with 2-pixel
halos. Inputs are synthetic (seeded) and resident in HBM before the timed region. One step = one
task (submit -> all 16 subtasks done). The timed output is compared byte for byte with the
oracle's output of the same image.

The default run (`--workload all`) also measures, each on its own configuration and under the
same clock, with its own roofline, CPU baseline and parity check:
  - matrixMultiply N=2048 float32, 16 subtasks of 512 x 512 (split-BF16 MFMA);
  - pageRank on a 10M-page power-law web graph, 20 iterations, 1M-page subtasks;
  - luDecomposition N=16384 (512 blocks, split-BF16 trailing update) + 8-RHS refined solve;
  - the other apps (matrixTranspose, prefixSum, sparseSolver, fft, pageRankInvertedIndex at the
    reference's default sizes), one record with a row per app;
  - the serving mix: seeded matrixMultiply / imageFiltering / pageRank requests served at
    saturation (all requests queued at once, `--serving-slots` clients) by the BATCHED and the
    PLAIN stub, then at two lower offered Poisson loads for latency.

Roofline: for each row the dominant kernel's algorithmic work per launch (bytes or flop) over its
device time, measured with HIP events on the stream the kernel runs on (BATCHED flushes that
launch that kernel family alone).

CPU baseline: the same task on the pm:: runtime's CPU stubs (one thread per core of the box's CPU
share), which run the apps' CPU callbacks (restatements of the reference's CPU callbacks) over
the same subtask decomposition; luDecomposition is timed at N=4096.

`--workload apps` measures only the other apps (the fft row as the line's headline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload all|imagefilter|matmul|pagerank|lu|apps]
                       [--no-cpu-baseline] [--serving-requests N] [--serving-slots S]
--steps / --warmup apply to the headline (or the one workload named); defaults per workload:
imagefilter and matmul 200 / 3, pagerank 10 / 1, lu 3 / 1, apps 8 / 1.
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
BF16_PEAK_TFLOPS = 2500.0
SEED = 20240601
# The dominant kernel's duration comes from HIP events on its stream around every KT_EVERY-th
# flush of the timed region: each event pair costs the device a few microseconds, which timing
# every flush would add to the tasks being measured.
KT_EVERY = 4
DEFAULTS = {"imagefilter": (200, 3), "matmul": (200, 3), "pagerank": (10, 1), "lu": (3, 1), "apps": (8, 1)}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cores():
    """Host cores of this GPU's CPU share (the box sets OMP_NUM_THREADS to it)."""
    n = os.cpu_count() or 1
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    return max(1, min(n, share if share > 0 else 16))


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


def lat_stats(lat):
    return {"p50_ms": round(1e3 * percentile(lat, 50), 4), "p99_ms": round(1e3 * percentile(lat, 99), 4)}


def timed_batched(U, run, steps, warmup, family, every=KT_EVERY, warm=True):
    """Pre-warm + warm-up + `steps` timed tasks in the BATCHED stub with graphs; returns
    (elapsed s, latencies, kernel ms per one-family flush, flushes timed, work per flush)."""
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    n_warm = prewarm(run) if warm else 0
    for _ in range(warmup):
        run()
    U.device_sync()
    U.set_kernel_timing(True, every=every)
    U.kernel_time(reset=True)
    elapsed, lat = time_tasks(run, steps)
    U.device_sync()
    ms, fl, work = U.kernel_time_family(family, reset=True)
    U.set_kernel_timing(False)
    return elapsed, lat, ms / max(1, fl), fl, work / max(1, fl), n_warm


def comparators(U, run, steps):
    out = {}
    for name, mode in (("plain", U.PM_HIP_MODE_PLAIN), ("resident", U.PM_HIP_MODE_RESIDENT)):
        U.set_hip_mode(mode)
        run()
        e, l = time_tasks(run, steps)
        out[name] = {"tasks_per_s": round(steps / e, 3), **lat_stats(l)}
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    return out


# ---------------------------------------------------------------- imageFiltering (headline)

def bench_imagefilter(args, U, steps, warmup, cores):
    H = W = 8192
    r, shift, tile = 2, 4, 2048
    rng = np.random.default_rng(SEED)
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    weights = rng.integers(-4, 5, (2 * r + 1) ** 2).astype(np.int8)
    f = U.ImageFilter(W, H, weights, r, mode=1, shift=shift, tile=tile)
    f.set_input(img)
    algo_bytes = 2 * W * H * 3
    log("imageFiltering 8192^2: batched")
    run = lambda: f.run(U.DEVICE_GPU)
    elapsed, lat, kernel_ms, flushes, work, warm = timed_batched(U, run, steps, warmup, U.FAMILY_FILTER)
    got = f.output()
    st = U.stats()
    comps = comparators(U, run, max(2, min(steps, args.comparator_steps)))

    # the same task alternating over two images (inputs and outputs 805 MB, three times the
    # 256 MB last-level cache): the kernel time is not a cache artefact
    img2 = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    f2 = U.ImageFilter(W, H, weights, r, mode=1, shift=shift, tile=tile)
    f2.set_input(img2)
    flip = [0]

    def run_alt():
        flip[0] ^= 1
        return (f if flip[0] else f2).run(U.DEVICE_GPU)
    e2, l2, k2, fl2, _, _ = timed_batched(U, run_alt, steps, warmup, U.FAMILY_FILTER, warm=False)
    comps["batched_two_images_alternating"] = {"tasks_per_s": round(steps / e2, 3), **lat_stats(l2),
                                               "kernel_ms": round(k2, 4), "timed_flushes": fl2}
    f2.close()

    cpu, parity = None, {"checked": False}
    if not args.no_cpu_baseline:
        cpu, parity = cpu_leg_imagefilter(U, f, img, weights, r, shift, tile, got, cores)
    f.close()
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    return {
        "metric": "tasks/s (served+batched, imageFiltering 8192^2 task)",
        "value": round(steps / elapsed, 3), "unit": "tasks/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 4), "dtype": "u8",
        "config": {"workload": "imageFiltering 8192x8192 RGB u8, 5x5 int weights in [-4,4], >>4 + clamp, "
                               "16 subtasks of 2048x2048 tiles with 2-px halo",
                   "hip_mode": "batched (one-family flushes launch directly)", "tile": tile, "radius": r},
        **lat_stats(lat),
        "bit_exact": parity.get("bit_exact"),
        "parity": parity,
        "per_subtask_launch": comps,
        "speedup_vs_plain": round((steps / elapsed) / comps["plain"]["tasks_per_s"], 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("imagefilter_8192"),
                     "kernel": "k_imagefilter_persistent", "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes,
                     "algorithmic_bytes_per_launch": algo_bytes, "library_counted_bytes_per_launch": work},
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
        "stub_stats": {k: st[k] for k in ("subtasks", "batches", "h2d_bytes", "d2h_bytes", "persistent_launches")},
    }


def cpu_leg_imagefilter(U, f, img, weights, r, shift, tile, got, cores):
    """CPU baseline: the task on the pm:: CPU stubs. Checker: the oracle's output of the same
    image (the C restatement of the reference callback), compared byte for byte with the GPU's."""
    from oracle import oracle_lib as O
    log("imageFiltering 8192^2: CPU stubs and oracle")
    t0 = time.perf_counter()
    f.run(U.DEVICE_CPU)
    dt = time.perf_counter() - t0
    cpu_out = f.output()
    H, W, _ = img.shape
    ntiles = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    want = np.zeros_like(img)
    O.imagefilter_tiled(img, weights, r, 1, shift, tile, 0, ntiles, cores, out=want)
    bad = int(np.count_nonzero(got != want))
    parity = {"checked": True, "bit_exact": bad == 0, "mismatched_bytes": bad, "against": "oracle (C restatement)",
              "cpu_stubs_bit_exact": bool(np.array_equal(cpu_out, want))}
    cpu = {"value": round(1.0 / dt, 4), "unit": "tasks/s", "cores": cores, "kind": "port",
           "sample": f"one full task ({ntiles} tiles of {tile}x{tile}) on the pm:: CPU stubs, {cores} threads, "
                     f"{dt:.2f} s"}
    return cpu, parity


# ---------------------------------------------------------------- matrixMultiply

def bench_matmul(args, U, steps, warmup, cores):
    """matrixMultiply N=2048 float32, A and B uniform in [-1, 1) from the seed, 16 subtasks of
    512 x 512 output blocks; split-BF16 MFMA products (3 per fp32 product) with fp32 accumulation.
    Roofline: MFMA-issued bf16 flop (3 * 2 N^3) per batched flush over the flush's device time."""
    n, bs = 2048, 512
    rng = np.random.default_rng(SEED)
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    B = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    mm.set_inputs(A, B)
    log("matrixMultiply 2048: batched")
    run = lambda: mm.run(U.DEVICE_GPU)
    elapsed, lat, kernel_ms, flushes, issued, warm = timed_batched(U, run, steps, warmup, U.FAMILY_GEMM)
    Cg = mm.output()
    comps = comparators(U, run, max(2, min(steps, args.comparator_steps)))
    C64 = A.astype(np.float64) @ B.astype(np.float64)
    err = float(np.linalg.norm(Cg - C64) / np.linalg.norm(C64))
    comp = float(np.abs(Cg - C64).max() / (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)).max())
    parity = {"checked": True, "fp64_normwise_rel_err": err, "max_componentwise_err_over_absAabsB": comp,
              "tolerance": 1e-5, "pass": err <= 1e-5, "against": "FP64 product"}
    cpu = None
    if not args.no_cpu_baseline:
        log("matrixMultiply 2048: CPU stubs")
        t0 = time.perf_counter()
        mm.run(U.DEVICE_CPU)
        dt = time.perf_counter() - t0
        ec = float(np.linalg.norm(mm.output() - C64) / np.linalg.norm(C64))
        cpu = {"value": round(1.0 / dt, 4), "unit": "tasks/s", "cores": cores, "kind": "port",
               "sample": f"one full task (16 subtasks, plain i-j-k float loop per 512x512 block) on the pm:: CPU "
                         f"stubs, {cores} threads, {dt:.2f} s; FP64 normwise error {ec:.1e}"}
    mm.close()
    flop = 2.0 * n ** 3
    achieved = issued / (kernel_ms * 1e-3) / 1e12
    return {
        "metric": "tasks/s (served+batched, matrixMultiply N=2048 fp32 task)",
        "value": round(steps / elapsed, 3), "unit": "tasks/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 4), "dtype": "f32 (split-bf16 x3 MFMA, fp32 accumulate)",
        "config": {"workload": "matrixMultiply N=2048 float32, 16 subtasks of 512x512", "hip_mode": "batched"},
        **lat_stats(lat), "parity": parity, "per_subtask_launch": comps,
        "fp32_equivalent_tflops": round(flop / (kernel_ms * 1e-3) / 1e12, 1),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": traffic_from_profiles("matmul_2048"),
                     "kernel": "k_gemm_persistent", "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes,
                     "algorithmic_flop_per_launch": issued},
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- pageRank

def bench_pagerank(args, U, steps, warmup, cores):
    """pageRank on a 10M-page synthetic power-law web graph (out-degree P(k) ~ (k + k0)^-2 on
    [1, 100], mean 10, 5 % dangling, preferential targets), 20 iterations, 1M-page subtasks,
    over the binned graph layout. One step = one pageRank computation (init task + 20 iteration
    tasks). Roofline: per iteration 12 N + 4 E bytes (row pointers, degree, rank in / out, in-link
    ids: the algorithmic bytes of a pull) over one iteration's flush time."""
    pages, per, iters = 10_000_000, 1_000_000, 20
    log("pageRank 10M: graph")
    rp, col = U.webgen(pages, SEED)
    edges = int(col.size)
    U.pr_set_layout("binned")
    pr = U.PageRank(rp, col, per)
    if pr.layout != "binned":
        raise RuntimeError("bench: the 10M-page graph did not take the binned layout")
    log("pageRank 10M: batched")
    run = lambda: pr.run(iters, U.DEVICE_GPU)
    elapsed, lat, kernel_ms, flushes, _, warm = timed_batched(U, run, steps, warmup, U.FAMILY_PAGERANK_BINNED)
    y = pr.ranks()
    comps = {}
    U.set_hip_mode(U.PM_HIP_MODE_PLAIN)
    e, l = time_tasks(run, 1)
    comps["plain"] = {"tasks_per_s": round(1 / e, 4), "p50_ms": round(1e3 * l[0], 1)}
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    cpu, parity = None, {"checked": False}
    if not args.no_cpu_baseline:
        log("pageRank 10M: CPU stubs and oracle")
        t0 = time.perf_counter()
        pr.run(iters, U.DEVICE_CPU)
        dt = time.perf_counter() - t0
        yc = pr.ranks()
        from oracle import oracle_lib as O
        yo = O.pagerank(rp, col, iters, per)
        err = float(np.abs(y.astype(np.float64) - yo).sum() / np.abs(yo).sum())
        parity = {"checked": True, "rel_l1_vs_oracle": err, "tolerance": 1e-5, "pass": err <= 1e-5,
                  "against": "oracle (push per subtask + REDUCE_ADD, fp32)",
                  "cpu_stubs_rel_l1_vs_oracle": float(np.abs(yc.astype(np.float64) - yo).sum() / np.abs(yo).sum())}
        cpu = {"value": round(1.0 / dt, 4), "unit": "tasks/s", "cores": cores, "kind": "port",
               "sample": f"one full task (20 iterations x 10 subtasks) on the pm:: CPU stubs, {cores} threads, "
                         f"{dt:.2f} s"}
    pr.close()
    log("pageRank 10M: source-blocked and CSR layouts, push decomposition comparators")
    for name, layout, fam in (("blocked_layout", "blocked", U.FAMILY_PAGERANK_BLOCKED),
                              ("csr_layout", "csr", U.FAMILY_PAGERANK)):
        U.pr_set_layout(layout)
        prc = U.PageRank(rp, col, per)
        U.pr_set_layout("binned")
        e, l, cms, cfl, _, _ = timed_batched(U, lambda: prc.run(iters, U.DEVICE_GPU), 1, 1, fam,
                                             every=1, warm=False)
        yl = prc.ranks()
        comps[name] = {"tasks_per_s": round(1 / e, 4), "kernel_ms_per_iteration": round(cms, 4),
                       "rel_l1_vs_binned": float(np.abs(yl.astype(np.float64) - y).sum() / np.abs(y).sum())}
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
    return {
        "metric": "tasks/s (served+batched, pageRank 10M pages x 20 iterations)",
        "value": round(steps / elapsed, 3), "unit": "tasks/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 3), "dtype": "f32",
        "config": {"workload": f"pageRank {pages} pages, {edges} links, 20 iterations, 1M-page subtasks",
                   "hip_mode": "batched"},
        **lat_stats(lat), "parity": parity, "per_subtask_launch": comps,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("pagerank_10m"),
                     "kernel": "k_pagerank_binned_persistent (one iteration)",
                     "kernel_ms": round(kernel_ms, 4), "timed_flushes": flushes,
                     "algorithmic_bytes_per_launch": it_bytes},
        "prewarm": {"seconds": PREWARM_S, "tasks": warm},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- luDecomposition

def _lu_system(rng, n, nrhs):
    M = rng.random((n, n), dtype=np.float32)
    M[np.diag_indices(n)] += np.float32(n)
    x_true = rng.uniform(-1, 1, (n, nrhs))
    b = np.asfortranarray(M.T.astype(np.float64) @ x_true)
    return M, x_true, b


def bench_lu(args, U, steps, warmup, cores):
    """luDecomposition N=16384 float32, A = U[0,1) + N I (column-major), 512 x 512 blocks, the
    trailing update in split-BF16 MFMA, then A x = b for 8 seeded right-hand sides with
    FP64-residual iterative refinement. One step = factorisation + refined solve, with the
    matrix resident in HBM. Errors are against the FP64 truth (b = A x_true in FP64). Roofline:
    the stage-3 trailing-update GEMM (k_gemm_persistent), its flop issued over its flushes'
    device time; the plain FP32 path (VALU update) is compared at N=4096."""
    n, bs, nrhs = 16384, 512, 8
    rng = np.random.default_rng(SEED)
    log("luDecomposition 16384: matrix")
    M, x_true, b = _lu_system(rng, n, nrhs)
    A = M.T                                   # column-major view of a U[0,1) + N I matrix
    lu = U.LU(n, bs, 0)
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    fact, solve, infos = [], [], []
    gemm = [0.0, 0, 0.0]
    trsm = [0.0, 0]
    for step in range(warmup + steps):
        lu.set_matrix(A)
        lu.prefetch()
        U.device_sync()
        timed = step >= warmup
        if timed:
            U.set_kernel_timing(True, every=1)
            U.kernel_time(reset=True)
        t0 = time.perf_counter()
        lu.factor(U.DEVICE_GPU)
        t1 = time.perf_counter()
        if timed:
            ms, fl, w = U.kernel_time_family(U.FAMILY_GEMM, reset=True)
            gemm[0] += ms; gemm[1] += fl; gemm[2] += w
            ms, fl, _ = U.kernel_time_family(U.FAMILY_TRSM, reset=True)
            trsm[0] += ms; trsm[1] += fl
            U.set_kernel_timing(False)
        x, info = lu.solve(b, 5, U.DEVICE_GPU)
        t2 = time.perf_counter()
        if timed:
            fact.append(t1 - t0)
            solve.append(t2 - t1)
            infos.append(info)
        log(f"lu: step {step + 1}/{warmup + steps} factor {t1 - t0:.4f} s solve {t2 - t1:.4f} s")
    fwd = float(np.abs(x - x_true).max() / np.abs(x_true).max())
    lu.close()
    del M, A
    factor_s, solve_s = float(np.median(fact)), float(np.median(solve))
    flop = 2.0 / 3.0 * n ** 3
    trailing_issued = 3 * 2.0 * sum((n - (k + 1) * bs) ** 2 * bs for k in range(n // bs - 1))
    n2 = 4096
    M2, x2, b2 = _lu_system(rng, n2, nrhs)
    small = {}
    for name, prec in (("split_bf16", 0), ("plain_fp32", 2)):
        log(f"lu: N={n2} comparison, {name}")
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
        log(f"lu: CPU stubs at N={n2}")
        l3 = U.LU(n2, bs, 0)
        l3.set_matrix(M2.T)
        t0 = time.perf_counter()
        l3.factor(U.DEVICE_CPU)
        t1 = time.perf_counter()
        y3, inf3 = l3.solve(b2, 5, U.DEVICE_CPU)
        t2 = time.perf_counter()
        l3.close()
        cpu = {"value": round(1.0 / (t2 - t0), 4), "unit": "tasks/s", "cores": cores, "kind": "port",
               "sample": f"one task at N={n2} (not 16384): factorisation {t1 - t0:.2f} s + refined solve "
                         f"{t2 - t1:.2f} s on the pm:: CPU stubs, {cores} threads; solve backward error "
                         f"{inf3['backward_error']:.1e}"}
    bwd = infos[-1]["backward_error"]
    parity = {"checked": True, "solve_backward_err": bwd, "solve_backward_tolerance": 2.0 ** -24,
              "forward_err": fwd, "pass": bool(bwd <= 2.0 ** -24),
              "against": "FP64 truth (b = A x_true in FP64); N=4096 factor errors in n4096_comparison"}
    gemm_ms, gemm_fl, gemm_w = gemm[0] / steps, gemm[1] / steps, gemm[2] / steps
    achieved = gemm_w / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    return {
        "metric": "tasks/s (luDecomposition N=16384 factorisation + 8-RHS refined solve)",
        "value": round(1.0 / (factor_s + solve_s), 4), "unit": "tasks/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * (factor_s + solve_s), 2),
        "dtype": "f32 (split-bf16 x3 MFMA update) + f64 residual",
        "config": {"workload": "luDecomposition N=16384 fp32 U[0,1)+N*I, 512 blocks, 8 RHS refinement",
                   "hip_mode": "batched"},
        "factor_s": round(factor_s, 4), "solve_s": round(solve_s, 4),
        "factor_fp32_equivalent_tflops": round(flop / factor_s / 1e12, 2),
        "factor_mfma_issued_tflops": round(trailing_issued / factor_s / 1e12, 1),
        "refine_steps": infos[-1]["steps"], "parity": parity, "n4096_comparison": small,
        "stage2_trsm": {"kernel": "k_trsm_persistent", "ms_per_factorisation": round(trsm[0] / steps, 3),
                        "flushes": trsm[1] // max(1, steps)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": traffic_from_profiles("lu_16384_stage3"),
                     "kernel": "k_gemm_persistent (stage-3 trailing update)",
                     "kernel_ms": round(gemm_ms / max(1, gemm_fl), 4), "timed_flushes": gemm_fl,
                     "kernel_ms_per_factorisation": round(gemm_ms, 3),
                     "algorithmic_flop_per_launch": round(gemm_w / max(1, gemm_fl)),
                     "algorithmic_flop_per_factorisation": gemm_w},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- serving mix

def bench_serving(args, U):
    """Serving mix on one GPU: seeded matrixMultiply (N log-uniform in 256..4096, float32,
    split-BF16 MFMA), imageFiltering (RGB u8 side uniform in 512..4096, 5x5, >>4 + clamp) and
    pageRank (100k..2M pages, 20 iterations) requests over resident datasets, each request's
    output returned to the host and hashed. Saturation: all requests queued at t = 0 and served by
    `slots` clients, on the BATCHED stub (continuous batching into persistent kernels) and on the
    PLAIN per-subtask-launch stub; equal accuracy = every request's output hash is the same on
    both. Latency: Poisson arrivals at 25 % and 50 % of the BATCHED capacity."""
    S = U.Server()
    t0 = time.perf_counter()
    mm = [256, 512, 768, 1024, 1536, 2048, 3072, 4096]
    cat = {"mm": [(n, S.add_mm(n, 512 if n % 512 == 0 else 256, 0, seed=100 + i)) for i, n in enumerate(mm)],
           "if": [(s_, S.add_if(s_, 2, 1, 4, 2048, seed=200 + i)) for i, s_ in enumerate(range(512, 4097, 512))],
           "pr": [(p_, S.add_pr(p_, 1_000_000, 20, seed=300 + i))
                  for i, p_ in enumerate([100_000, 500_000, 1_000_000, 2_000_000])]}
    setup_s = time.perf_counter() - t0
    n, slots = args.serving_requests, args.serving_slots
    ds, _ = U.serving_stream(SEED, n, 0.0, cat)
    kind_of = {d: k for k, entries in cat.items() for _, d in entries}
    kinds = np.array([kind_of[int(d)] for d in ds])
    out = {"requests": n, "slots": slots, "seed": SEED,
           "mix": {k: int((kinds == k).sum()) for k in ("mm", "if", "pr")},
           "results_to_host": True, "dataset_setup_s": round(setup_s, 2)}
    warm = np.array(sorted({int(d) for d in ds}), dtype=np.int32)

    def serve(mode, arrivals, label):
        U.set_hip_mode(mode, graphs=True)
        S.run(warm, np.zeros(len(warm)), 4, U.DEVICE_GPU)          # uploads, clones, graphs
        log(f"serving: {label}")
        lat, cs, wall = S.run(ds, arrivals, slots, U.DEVICE_GPU, checksums=True)
        res = {"tasks_per_s": round(n / wall, 2), "wall_s": round(wall, 3),
               "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
               "per_class": {k: {"p50_ms": round(percentile(lat[kinds == k], 50), 3),
                                 "p99_ms": round(percentile(lat[kinds == k], 99), 3)}
                             for k in ("mm", "if", "pr") if (kinds == k).any()}}
        return res, cs

    zero = np.zeros(n)
    out["saturation_batched"], cs_b = serve(U.PM_HIP_MODE_BATCHED, zero, "saturation, batched")
    out["saturation_plain"], cs_p = serve(U.PM_HIP_MODE_PLAIN, zero, "saturation, plain")
    out["equal_outputs_batched_vs_plain"] = bool(np.array_equal(cs_b, cs_p))
    cap = out["saturation_batched"]["tasks_per_s"]
    out["speedup_vs_plain"] = round(cap / out["saturation_plain"]["tasks_per_s"], 2)
    loads = []
    for frac in (0.25, 0.5):
        rate = frac * cap
        _, arr = U.serving_stream(SEED, n, rate, cat)
        res, cs = serve(U.PM_HIP_MODE_BATCHED, arr, f"offered {rate:.0f}/s, batched")
        res["offered_per_s"] = round(rate, 1)
        res["equal_outputs"] = bool(np.array_equal(cs, cs_b))
        loads.append(res)
    out["offered_load_batched"] = loads
    U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
    S.close()
    return out


# ---------------------------------------------------------------- the other apps (--workload apps)

def kernel_ms(launch, reps=20):
    """Average device time of `launch(stream)` over `reps` back-to-back calls, from HIP events on
    the stream the kernel runs on (after two untimed calls)."""
    from unicorn_amd import hip as H
    L = H.lib()
    st = H.Stream()
    e0, e1 = C_void(), C_void()
    H.check(L.uh_event_create(C_ref(e0)), "event")
    H.check(L.uh_event_create(C_ref(e1)), "event")
    for _ in range(2):
        launch(st.h)
    st.sync()
    H.check(L.uh_event_record(e0.value, st.h), "record")
    for _ in range(reps):
        launch(st.h)
    H.check(L.uh_event_record(e1.value, st.h), "record")
    H.check(L.uh_event_sync(e1.value), "event sync")
    ms = C_float()
    H.check(L.uh_event_elapsed_ms(C_ref(ms), e0.value, e1.value), "elapsed")
    L.uh_event_destroy(e0.value)
    L.uh_event_destroy(e1.value)
    return ms.value / reps


def C_void():
    import ctypes
    return ctypes.c_void_p()


def C_float():
    import ctypes
    return ctypes.c_float()


def C_ref(x):
    import ctypes
    return ctypes.byref(x)


def app_row(name, workload, run_gpu, steps, parity, cpu_run, cores, kernel, dtype):
    """tasks/s of one app task on the HIP stub (RESIDENT: per-subtask launches of the app's GPU
    callback on HBM mirrors), with the dominant kernel's roofline and the CPU-stub baseline."""
    import unicorn_amd as U
    U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
    run_gpu()
    elapsed, lat = time_tasks(run_gpu, steps)
    cpu = None
    if cpu_run is not None:
        t0 = time.perf_counter()
        cpu_run()
        dt = time.perf_counter() - t0
        cpu = {"value": round(1.0 / dt, 4), "unit": "tasks/s", "cores": cores, "kind": "port",
               "sample": f"one full task on the pm:: CPU stubs (the app's CPU callbacks), {cores} threads, {dt:.3f} s"}
    return {"metric": f"tasks/s ({name} task, HIP stub resident)", "value": round(steps / elapsed, 3),
            "unit": "tasks/s", "steps": steps, "ms_per_step": round(1e3 * elapsed / steps, 4), "dtype": dtype,
            "config": {"workload": workload, "hip_mode": "resident"}, **lat_stats(lat), "parity": parity,
            "roofline": kernel, "cpu_baseline": cpu}


def hbm_roofline(kernel_name, nbytes, ms, what, profile=None):
    """`profile` names the profiles/<profile>_traffic.json holding the kernel's measured DRAM bytes."""
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(profile) if profile else None,
            "kernel": kernel_name, "kernel_ms": round(ms, 5), "algorithmic_bytes_per_launch": int(nbytes),
            "per_launch": what}


def bench_apps(args, U, steps, warmup, cores):
    """The apps outside the headline configurations, each at the reference's default size:
    matrixTranspose 16384 x 4096 int32 (2048-square blocks), prefixSum of 10M int32 (1M per
    subtask), sparseSolver n = 1024 at density 0.001, fft 4096 x 4096 complex float (512 rows /
    columns per subtask), pageRankInvertedIndex over a 1M-page web dump (the reference's dump has
    100M pages; 250K pages per subtask, 2 iterations). One step = one app task (both tasks for fft,
    the inversion plus both iterations for pageRankInvertedIndex). The headline of the line is
    fft; the others are under rows."""
    from unicorn_amd import hip as H
    from oracle import oracle_lib as O
    L = H.lib()
    rows = {}
    rng = np.random.default_rng(SEED)
    use_cpu = not args.no_cpu_baseline

    # fft (headline of this line)
    log("apps: fft 4096^2")
    n = 4096
    x = (rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))).astype(np.complex64)
    f = U.Fft2D(n, n, 512)
    f.set_input(x)
    U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
    f.run(U.DEVICE_GPU)
    got = f.output()
    ref = O.fft2(x)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    parity = {"checked": True, "rel_l2_vs_fp64": err, "tolerance": 1e-5, "pass": err <= 1e-5,
              "against": "oracle 2-D DFT in double"}
    d_in, d_out = H.DeviceBuffer.from_array(x), H.DeviceBuffer(x.nbytes)
    ms_row = kernel_ms(lambda s: L.uh_fft_c2c_f32(d_in.ptr, n, 1, d_out.ptr, n, 1, n, 512, s))
    ms_col = kernel_ms(lambda s: L.uh_fft_c2c_f32(d_in.ptr, 1, n, d_out.ptr, 1, n, n, 512, s))
    d_in.free()
    d_out.free()
    kern = hbm_roofline("k_fft (columns)", 512 * n * 16, ms_col, "512 columns of 4096 points (NATURAL view)",
                        "fft_4096_columns")
    kern["rows_kernel"] = hbm_roofline("k_fft (rows)", 512 * n * 16, ms_row, "512 rows of 4096 points", "fft_4096_rows")
    head = app_row("fft 4096x4096", "fft 4096x4096 complex float, 512 rows / columns per subtask",
                   lambda: f.run(U.DEVICE_GPU), steps, parity, (lambda: f.run(U.DEVICE_CPU)) if use_cpu else None,
                   cores, kern, "c64 (f32 pairs)")
    f.close()

    log("apps: matrixTranspose 16384x4096")
    R, Cc = 1 << 14, 1 << 12
    M = rng.integers(-2**31, 2**31 - 1, (R, Cc), dtype=np.int64).astype(np.int32)
    mt = U.MatTranspose(R, Cc)
    mt.set_input(M)
    U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
    mt.run(U.DEVICE_GPU)
    ok = bool(np.array_equal(mt.output(), M.T))
    d_in, d_out = H.DeviceBuffer.from_array(M[:2048]), H.DeviceBuffer(2048 * 2048 * 4)
    ms = kernel_ms(lambda s: L.uh_transpose32(d_in.ptr, Cc, d_out.ptr, 2048, 2048, 2048, s))
    d_in.free()
    d_out.free()
    rows["matrixTranspose"] = app_row(
        "matrixTranspose 16384x4096", "matrixTranspose 16384x4096 int32 out of place, 2048-square blocks",
        lambda: mt.run(U.DEVICE_GPU), steps, {"checked": True, "bit_exact": ok, "against": "M.T"},
        (lambda: mt.run(U.DEVICE_CPU)) if use_cpu else None, cores,
        hbm_roofline("k_transpose32", 2 * 2048 * 2048 * 4, ms, "one 2048x2048 block", "transpose_2048"), "i32")
    mt.close()

    log("apps: prefixSum 10M")
    N = 10_000_000
    v = rng.integers(-1000, 1000, N).astype(np.int32)
    ps = U.PrefixSum(N, 1_000_000)
    ps.set_input(v)
    U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
    ps.run(U.DEVICE_GPU)
    ok = bool(np.array_equal(ps.output(), O.prefix_sum(v)))
    d_in, d_out = H.DeviceBuffer.from_array(v[:1_000_000]), H.DeviceBuffer(4_000_000)
    ms = kernel_ms(lambda s: L.uh_scan_i32(d_in.ptr, d_out.ptr, 1_000_000, s))
    d_in.free()
    d_out.free()
    rows["prefixSum"] = app_row(
        "prefixSum 10M", "prefixSum 10,000,000 int32, 1,000,000 per subtask (scan task, host aux scan, add task)",
        lambda: ps.run(U.DEVICE_GPU), steps, {"checked": True, "bit_exact": ok, "against": "oracle serialPrefixSum"},
        (lambda: ps.run(U.DEVICE_CPU)) if use_cpu else None, cores,
        hbm_roofline("k_scan_totals + k_scan_blocks", 8 * 1_000_000, ms, "one 1,000,000-element chunk (both passes)",
                     "prefix_sum_1m"),
        "i32")
    ps.close()

    log("apps: sparseSolver 1024")
    n = 1024
    k = max(1, int(n * n * 0.001))
    idx = np.unique(rng.choice(n * n, size=k, replace=False))
    r_, c_ = (idx // n).astype(np.int32), (idx % n).astype(np.int32)
    vals = rng.standard_normal(r_.size).astype(np.float32)
    B = rng.standard_normal((n, n)).astype(np.float32)
    sm = U.SparseMatMul(n, r_, c_, vals, B)
    U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
    sm.run(U.DEVICE_GPU)
    ok = bool(np.array_equal(sm.output(), O.spmm_coo(r_, c_, vals, B)))
    ent = np.zeros((r_.size, 3), np.uint32)
    ent[:, 0], ent[:, 1], ent[:, 2] = r_.view(np.uint32), c_.view(np.uint32), vals.view(np.uint32)
    d_e, d_b, d_c = H.DeviceBuffer.from_array(ent), H.DeviceBuffer.from_array(B), H.DeviceBuffer(n * n * 4)
    ms = kernel_ms(lambda s: L.uh_spmm_coo_f32(d_e.ptr, r_.size, 0, n, d_b.ptr, n, d_c.ptr, n, n, s))
    d_e.free(); d_b.free(); d_c.free()
    rows["sparseSolver"] = app_row(
        "sparseSolver 1024", f"sparseSolver n=1024, {r_.size} non-zeros (density 0.001), one 1024-square block",
        lambda: sm.run(U.DEVICE_GPU), steps, {"checked": True, "bit_exact": ok, "against": "oracle COO loop (float)"},
        (lambda: sm.run(U.DEVICE_CPU)) if use_cpu else None, cores,
        hbm_roofline("k_spmm_coo", 12 * r_.size + 4 * r_.size * n + 4 * n * n, ms,
                     "one 1024-row block: entries, the B rows they select, C", "sparse_1024"), "f32")
    sm.close()

    log("apps: pageRankInvertedIndex 1M pages")
    import tempfile
    pages, P = 1_000_000, 250_000
    rp, col = U.webgen(pages, SEED)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        U.webdump_write(tmp, rp, col, 100, 50_000)
        a = U.PageRankInvertedIndex(tmp, P)
        U.set_hip_mode(U.PM_HIP_MODE_RESIDENT)
        a.run(U.DEVICE_GPU, 2)
        e, counts = a.index()
        y = a.ranks()
        recs = O.prii_records(rp, col, 100)
        eo, co = O.prii_invert(recs, P)
        yo = O.prii_pagerank(eo, co, pages, P, 2).astype(np.float64)
        rel = float(np.abs(y - yo).sum() / np.abs(yo).sum())
        parity = {"checked": True, "index_bit_exact": bool(np.array_equal(e, eo)), "ranks_rel_l1_vs_oracle": rel,
                  "tolerance": 1e-5, "pass": bool(np.array_equal(e, eo)) and rel <= 1e-5,
                  "against": "oracle invertIndex_cpu + redistribution, pageRank_cpu"}
        sub = np.ascontiguousarray(recs[:P])
        links = int(sub[:, 0].sum())
        d_d, d_o, d_t = H.DeviceBuffer.from_array(sub), H.DeviceBuffer(P * 100 * 12), H.DeviceBuffer(4 * 4)
        ms_inv = kernel_ms(lambda s: L.uh_invert_index(d_d.ptr, P, 100, 0, pages, P, 4, d_o.ptr, d_t.ptr, s), reps=5)
        d_d.free(); d_o.free(); d_t.free()
        n0 = int(co[0])
        d_e, d_x, d_y = H.DeviceBuffer.from_array(eo[:n0]), H.DeviceBuffer.from_array(y), H.DeviceBuffer(P * 4)
        ms_rank = kernel_ms(lambda s: L.uh_prii_rank(d_e.ptr, n0, d_x.ptr, 1.0, 0.85, 0, d_y.ptr, P, pages, s))
        d_e.free(); d_x.free(); d_y.free()
        kern = hbm_roofline("uh_invert_index (4 kernels)", 4 * P + 4 * links + 12 * links, ms_inv,
                            f"one 250K-page subtask: counts and {links} links read, {links} entries written",
                            "prii_invert_250k")
        kern["rank_kernel"] = hbm_roofline("k_prii_rank", 12 * n0 + 4 * n0 + 4 * P, ms_rank,
                                           f"one subtask's {n0} inlinks: entries, source ranks, output ranks",
                                           "prii_rank_250k")
        rows["pageRankInvertedIndex"] = app_row(
            "pageRankInvertedIndex 1M pages", f"pageRankInvertedIndex {pages} pages ({col.size} links) from a web "
            f"dump, 250K pages per subtask, inversion + 2 iterations", lambda: a.run(U.DEVICE_GPU, 2), max(2, steps // 4),
            parity, (lambda: a.run(U.DEVICE_CPU, 2)) if use_cpu else None, cores, kern, "u32 index, f32 ranks")
        a.close()
    head.update({"warmup": 1, "rows": rows})
    return head


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


BENCHES = {"imagefilter": bench_imagefilter, "matmul": bench_matmul, "pagerank": bench_pagerank, "lu": bench_lu,
           "apps": bench_apps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="all", choices=["all"] + list(BENCHES))
    ap.add_argument("--comparator-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=300, help="0 skips the serving mix")
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run D2D copies of one image size (reference for HBM counters)")
    args = ap.parse_args()
    primary = "imagefilter" if args.workload == "all" else args.workload
    steps, warmup = DEFAULTS[primary]
    if args.steps is not None:
        steps = args.steps
    if args.warmup is not None:
        warmup = args.warmup
    if steps < 1 or warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only (no collectives or multi-GPU path); "
                                   "run with --gpus 1"}), flush=True)
        return 2
    import unicorn_amd as U
    cores = host_cores()
    U.init(cores)
    if not U.has_gpu():
        print(json.dumps({"error": "no HIP device visible"}), flush=True)
        return 1
    if args.copy_calibration:
        copy_calibration(8192 * 8192 * 3)
    res = BENCHES[primary](args, U, steps, warmup, cores)
    line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": res.pop("steps"), "warmup": res.pop("warmup"), "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic (seeded)", "config": res.pop("config")}
    line.update(res)
    line["host_cores"] = cores
    if args.workload == "all":
        rows = {}
        for name in ("matmul", "pagerank", "lu", "apps"):
            s, w = DEFAULTS[name]
            rows[name] = BENCHES[name](args, U, s, w, cores)
        line["rows"] = rows
        if args.serving_requests > 0:
            line["serving"] = bench_serving(args, U)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
