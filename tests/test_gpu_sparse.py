"""sparseSolver through the HIP stub (uh_spmm_coo_f32 per block row of C), bit-exact against the
CPU callback's order of operations (the oracle's float form) in every stub mode and on both stub
kinds, including block rows without entries and the reference's default size (n = 1024,
density 0.001, sparseSolver.h:38-41)."""
import numpy as np
import pytest

from oracle import oracle_lib as O
from test_sparse import make, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("n,density", [(64, 0.05), (1024, 0.001), (96, 0.2), (4096, 0.0000002)])
def test_sparse_gpu(gpu_runtime, hip_mode, n, density):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rows, cols, vals, B = make(n, density, n + hip_mode)
    got = run(U, n, rows, cols, vals, B, U.DEVICE_GPU)
    np.testing.assert_array_equal(got, O.spmm_coo(rows, cols, vals, B))


def test_sparse_both_devices(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(0)
    rows, cols, vals, B = make(4096, 0.0001, 9)
    got = run(U, 4096, rows, cols, vals, B, U.DEVICE_BOTH)
    np.testing.assert_array_equal(got, O.spmm_coo(rows, cols, vals, B))
