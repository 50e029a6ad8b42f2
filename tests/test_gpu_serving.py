"""Served mode on the GPU: a seeded Poisson stream of mixed matrixMultiply / imageFiltering /
pageRank requests over resident datasets.

Every kernel of the path is deterministic (fixed summation orders, no atomics in the float
paths), so each request's output must be bit-identical whether it is served alone (one slot,
requests back to back, PLAIN per-subtask launches) or in the mix (concurrent requests whose
subtasks share persistent-kernel batches in the BATCHED stub, with hipGraph replay).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _catalogue(S):
    return {"mm": [(256, S.add_mm(256, 256, seed=11)), (512, S.add_mm(512, 512, seed=12)),
                   (1024, S.add_mm(1024, 512, seed=13))],
            "if": [(512, S.add_if(512, seed=21)), (1024, S.add_if(1024, seed=22)),
                   (2048, S.add_if(2048, seed=23))],
            "pr": [(100_000, S.add_pr(100_000, 60_000, 5, seed=31)),
                   (200_000, S.add_pr(200_000, 60_000, 5, seed=32))]}


def test_serving_alone_equals_mix(gpu_runtime):
    U = gpu_runtime
    S = U.Server()
    try:
        cat = _catalogue(S)
        ds, arr = U.serving_stream(2024, 60, 4000.0, cat)
        U.set_hip_mode(U.PM_HIP_MODE_PLAIN)
        _, plain_alone, _ = S.run(ds, np.zeros(len(ds)), 1, U.DEVICE_GPU, checksums=True)
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=False)
        _, batched_alone, _ = S.run(ds, np.zeros(len(ds)), 1, U.DEVICE_GPU, checksums=True)
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
        lat, mix, wall = S.run(ds, arr, 8, U.DEVICE_GPU, checksums=True)
    finally:
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
        S.close()
    assert np.array_equal(plain_alone, batched_alone)
    assert np.array_equal(batched_alone, mix), f"{np.count_nonzero(batched_alone != mix)} requests differ"
    assert np.all(lat > 0) and wall > 0


def test_serving_hbm_budget_admission(gpu_runtime):
    """With a budget that holds only some datasets at once, requests still all complete (idle
    mirrors are evicted, blocked requests wait) and give the same outputs."""
    U = gpu_runtime
    S = U.Server()
    try:
        cat = _catalogue(S)
        ds, arr = U.serving_stream(77, 40, 4000.0, cat)
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
        _, want, _ = S.run(ds, np.zeros(len(ds)), 1, U.DEVICE_GPU, checksums=True)
        U.set_hbm_budget(48 << 20)     # the largest request needs 25 MB; all datasets ~70 MB
        _, got, _ = S.run(ds, arr, 8, U.DEVICE_GPU, checksums=True)
        st = U.stats()
    finally:
        U.set_hbm_budget(0)
        S.close()
    assert np.array_equal(want, got)
    assert st["evictions"] > 0
