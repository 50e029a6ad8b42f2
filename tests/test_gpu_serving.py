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


def _rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def test_serving_mix_at_config_sizes_matches_oracle(gpu_runtime):
    """The serving mix at the sizes its configuration names: matrixMultiply N up to 4096,
    imageFiltering sides up to 4096 (and a ragged 3000), pageRank up to 2M pages with 1M-page
    subtasks and 20 iterations. Every request's output is checked: its hash must equal the hash
    of an output that was itself checked against the oracle (imageFiltering: the oracle's own
    output, bit-exact) or FP64 (matrixMultiply normwise <= 1e-5 vs an FP64 product; pageRank
    relative L1 <= 1e-5 vs the oracle, and the FP64 recurrence reported). The checked outputs
    of matrixMultiply and pageRank come from the same dataset run alone on the PLAIN stub, so
    the check also holds each mixed request equal to its alone result."""
    from oracle import oracle_lib as O
    U = gpu_runtime
    S = U.Server()
    pr_seed = {100_000: 61, 2_000_000: 62}
    try:
        cat = {"mm": [(256, S.add_mm(256, 256, seed=41)), (1024, S.add_mm(1024, 512, seed=42)),
                      (4096, S.add_mm(4096, 512, seed=43))],
               "if": [(512, S.add_if(512, 2, 1, 4, 2048, seed=51)), (3000, S.add_if(3000, 2, 1, 4, 2048, seed=52)),
                      (4096, S.add_if(4096, 2, 1, 4, 2048, seed=53))],
               "pr": [(p, S.add_pr(p, 1_000_000, 20, seed=sd)) for p, sd in pr_seed.items()]}
        ds, arr = U.serving_stream(4242, 48, 3000.0, cat)
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
        print(f"serving mix: {len(ds)} requests over {len(set(ds.tolist()))} datasets", flush=True)
        lat, got, wall = S.run(ds, arr, 8, U.DEVICE_GPU, checksums=True)
        U.set_hip_mode(U.PM_HIP_MODE_PLAIN)
        expected = {}
        for kind, entries in cat.items():
            for size, d in entries:
                if not np.any(ds == d):
                    continue
                print(f"serving mix: checking {kind} {size}", flush=True)
                if kind == "if":
                    img, w, r, mode, shift = S.if_input(d)
                    want = O.imagefilter(img, w, r, mode, shift)
                    expected[d] = U.output_hash(want)
                elif kind == "mm":
                    A, B = S.mm_inputs(d)
                    Cm = S.run_alone(d, U.DEVICE_GPU)
                    C64 = A.astype(np.float64) @ B.astype(np.float64)
                    err = float(np.linalg.norm(Cm - C64) / np.linalg.norm(C64))
                    print(f"  mm {size}: FP64 normwise error {err:.2e}", flush=True)
                    assert err <= 1e-5
                    expected[d] = U.output_hash(Cm)
                else:
                    rp, col = U.webgen(size, pr_seed[size])     # the generator add_pr used
                    y = S.run_alone(d, U.DEVICE_GPU, 20)
                    e_or = _rel_l1(y, O.pagerank(rp, col, 20, 1_000_000))
                    e64 = _rel_l1(y, O.pagerank_f64(rp, col, 20))
                    print(f"  pr {size}: relative L1 {e_or:.2e} vs oracle, {e64:.2e} vs FP64", flush=True)
                    assert e_or <= 1e-5 and e64 <= 1e-5
                    expected[d] = U.output_hash(y)
    finally:
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED, graphs=True)
        S.close()
    assert set(expected) == set(ds.tolist())
    bad = [i for i in range(len(ds)) if int(got[i]) != expected[int(ds[i])]]
    assert not bad, f"{len(bad)} of {len(ds)} requests differ from their checked output (first: request {bad[0]})"
    assert np.all(lat > 0) and wall > 0
