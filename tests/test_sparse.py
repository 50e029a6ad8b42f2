"""sparseSolver on the CPU stubs (no GPU needed): C = A B with A sparse in COO order and B dense,
one subtask per block of C (reference sparseSolver.cpp). Entries of each output accumulate in COO
order with each product and sum rounded, as the oracle's float form: bit-exact; and within the
reference's 1e-5 relative check of the FP64 product (sparseSolver.cpp:540-551)."""
import numpy as np
import pytest

from oracle import oracle_lib as O


def make(n, density, seed, ints=False):
    rng = np.random.default_rng(seed)
    k = max(1, int(n * n * density))
    idx = np.unique(rng.choice(n * n, size=k, replace=False))
    rows, cols = (idx // n).astype(np.int32), (idx % n).astype(np.int32)
    if ints:
        vals = rng.integers(-9, 10, rows.size).astype(np.float32)
        B = rng.integers(-9, 10, (n, n)).astype(np.float32)
    else:
        vals = rng.standard_normal(rows.size).astype(np.float32)
        B = rng.standard_normal((n, n)).astype(np.float32)
    return rows, cols, vals, B


def run(U, n, rows, cols, vals, B, device):
    m = U.SparseMatMul(n, rows, cols, vals, B)
    try:
        m.run(device)
        return m.output()
    finally:
        m.close()


@pytest.mark.parametrize("n,density", [(64, 0.05), (256, 0.001), (96, 0.2)])
def test_sparse_cpu(runtime, n, density):
    U = runtime
    rows, cols, vals, B = make(n, density, n)
    got = run(U, n, rows, cols, vals, B, U.DEVICE_CPU)
    np.testing.assert_array_equal(got, O.spmm_coo(rows, cols, vals, B))
    ref = O.spmm_coo(rows, cols, vals, B, f64=True)
    assert np.linalg.norm(got - ref) <= 1e-5 * max(np.linalg.norm(ref), 1e-30)


def test_sparse_cpu_empty_block_rows(runtime):
    U = runtime
    n = 4096                                  # 2 x 2 blocks of 2048; entries only in the first block row
    rows = np.array([0, 5, 2047], dtype=np.int32)
    cols = np.array([3, 4095, 0], dtype=np.int32)
    vals = np.array([2.0, -1.0, 0.5], dtype=np.float32)
    B = np.random.default_rng(1).integers(-4, 5, (n, n)).astype(np.float32)
    got = run(U, n, rows, cols, vals, B, U.DEVICE_CPU)
    np.testing.assert_array_equal(got, O.spmm_coo(rows, cols, vals, B))


def test_sparse_rejects_unsorted(runtime):
    with pytest.raises(Exception):
        runtime.SparseMatMul(8, np.array([1, 0], np.int32), np.array([0, 0], np.int32),
                             np.ones(2, np.float32), np.ones((8, 8), np.float32))
