"""Host logic of the pm:: runtime on the CPU stubs (no GPU needed).

The same app drivers as the GPU tests run here with their _CPU callback registrations, so the
task submission, data distribution, subscription views (NATURAL and COMPACT), reductions and
completion paths of libpmlib.so are exercised against the oracle:
  - imageFiltering: bit-exact (reference imageFiltering.cpp:379-472 restated);
  - matrixMultiply: bit-exact on integer-valued floats (plain i-j-k, matrixMultiply.cpp:47-55);
  - pageRank: relative L1 <= 1e-6 against the oracle (summation order differs);
  - luDecomposition: factors within 1e-4 of the unblocked oracle, refined solve to 2^-24.
Error paths: a GPU-only task where no HIP device exists fails with pmNoCompatibleDevice (20).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

PM_NO_COMPATIBLE_DEVICE = 20


def test_imagefilter_cpu_stub_matches_oracle(runtime):
    U = runtime
    rng = np.random.default_rng(3)
    for (H, W, r, mode, tile) in [(130, 257, 2, 0, 64), (67, 45, 1, 1, 32), (20, 20, 3, 0, 7)]:
        img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        w = rng.integers(-128 if mode == 0 else -4, 128 if mode == 0 else 5, (2 * r + 1) ** 2).astype(np.int8)
        f = U.ImageFilter(W, H, w, r, mode=mode, shift=4, tile=tile)
        try:
            f.set_input(img)
            f.run(U.DEVICE_CPU)
            got = f.output()
        finally:
            f.close()
        assert np.array_equal(got, O.imagefilter(img, w, r, mode, 4)), (H, W, r, mode, tile)


def test_matmul_cpu_stub_integer_exact(runtime):
    U = runtime
    rng = np.random.default_rng(5)
    n, bs = 256, 64
    A = rng.integers(-8, 9, (n, n)).astype(np.float32)
    B = rng.integers(-8, 9, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    try:
        mm.set_inputs(A, B)
        mm.run(U.DEVICE_CPU)
        C = mm.output()
    finally:
        mm.close()
    assert np.array_equal(C, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))


def test_pagerank_cpu_stub_matches_oracle(runtime):
    U = runtime
    rp, col = U.webgen(20_000, 9)
    pr = U.PageRank(rp, col, 6_000)
    try:
        pr.run(5, U.DEVICE_CPU)
        y = pr.ranks()
    finally:
        pr.close()
    want = O.pagerank(rp, col, 5, 6_000)
    rel = np.abs(y.astype(np.float64) - want).sum() / np.abs(want).sum()
    assert rel <= 1e-6


def test_lu_cpu_stub_factor_and_refine(runtime):
    U = runtime
    n, B = 256, 64
    rng = np.random.default_rng(11)
    A = np.asfortranarray((rng.random((n, n)) + n * np.eye(n)).astype(np.float32))
    lu = U.LU(n, B, 2)
    try:
        lu.set_matrix(A)
        lu.factor(U.DEVICE_CPU)
        F = lu.factors()
        x_true = rng.uniform(-1, 1, (n, 2))
        b = A.astype(np.float64) @ x_true
        x, info = lu.solve(np.asfortranarray(b), 5, U.DEVICE_CPU)
    finally:
        lu.close()
    Fo = O.lu(A)
    assert np.abs(F - Fo).max() <= 1e-4 * np.abs(Fo).max()
    assert info["backward_error"] <= 2.0 ** -24
    assert np.abs(x - x_true).max() <= 1e-6


def test_gpu_only_task_without_device_fails_cleanly(runtime):
    U = runtime
    if U.has_gpu():
        pytest.skip("a HIP device is present; this checks the no-device error path")
    f = U.ImageFilter(16, 16, np.ones(9, np.int8), 1, mode=0, shift=0, tile=8)
    try:
        f.set_input(np.zeros((16, 16, 3), np.uint8))
        with pytest.raises(RuntimeError, match=f"status {PM_NO_COMPATIBLE_DEVICE}"):
            f.run(U.DEVICE_GPU)
    finally:
        f.close()


def test_serving_alone_equals_mix_cpu(runtime):
    """Each request's output is the same served alone (one slot, back to back) and in a Poisson
    mix of concurrent requests over shared resident inputs (CPU stubs)."""
    U = runtime
    S = U.Server()
    try:
        cat = {"mm": [(64, S.add_mm(64, 32, seed=1)), (128, S.add_mm(128, 64, seed=2))],
               "if": [(64, S.add_if(64, tile=32, seed=3)), (130, S.add_if(130, tile=64, seed=4))],
               "pr": [(2000, S.add_pr(2000, 700, 4, seed=5)), (5000, S.add_pr(5000, 2000, 4, seed=6))]}
        ds, arr = U.serving_stream(7, 48, 3000.0, cat)
        _, alone, _ = S.run(ds, np.zeros(len(ds)), 1, U.DEVICE_CPU, checksums=True)
        lat, mix, wall = S.run(ds, arr, 6, U.DEVICE_CPU, checksums=True)
    finally:
        S.close()
    assert np.array_equal(alone, mix)
    assert np.all(lat >= 0) and wall > 0
    assert len(set(ds.tolist())) >= 4          # the stream mixes classes and sizes
