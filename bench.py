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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
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


def percentile(v, q):
    return float(np.percentile(np.asarray(v, dtype=np.float64), q)) if v else None


def time_tasks(run, steps):
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        lat.append(run())
    t1 = time.perf_counter()
    return t1 - t0, lat


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
    U.set_kernel_timing(True)
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
                "kernel": "k_imagefilter_persistent", "kernel_ms": round(kernel_ms, 4),
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
        "cpu_baseline": cpu,
        "stub_stats": {k: st[k] for k in ("subtasks", "batches", "h2d_bytes", "d2h_bytes", "persistent_launches")},
    }


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
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="imagefilter", choices=["imagefilter"])
    ap.add_argument("--comparator-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}), flush=True)
        return 2
    import unicorn_amd as U
    U.init(0)
    if not U.has_gpu():
        print(json.dumps({"error": "no HIP device visible"}), flush=True)
        return 1
    res = bench_imagefilter(args, U)
    line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic (seeded)", "config": res.pop("config")}
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
