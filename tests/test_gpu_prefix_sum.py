"""prefixSum through the HIP stub (uh_scan_i32 per chunk, uh_add_scalar_i32 per later chunk),
bit-exact against the oracle, in every stub mode and on both stub kinds, up to the reference's
default size (10,000,000 elements in chunks of 1,000,000, prefixSum.h:40-41)."""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("n,chunk", [(10007, 1000), (4096 * 3 + 5, 4096 * 3 + 5), (1, 1), (300001, 65536)])
def test_prefix_sum_gpu(gpu_runtime, hip_mode, n, chunk):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    x = np.random.default_rng(n + hip_mode).integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32)
    ps = U.PrefixSum(n, chunk)
    try:
        ps.set_input(x)
        ps.run(U.DEVICE_GPU)
        got = ps.output()
    finally:
        ps.close()
    np.testing.assert_array_equal(got, O.prefix_sum(x))


def test_prefix_sum_reference_default_size(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(2)
    n = 10_000_000
    x = np.random.default_rng(3).integers(-1000, 1000, n).astype(np.int32)
    ps = U.PrefixSum(n)
    try:
        ps.set_input(x)
        ps.run(U.DEVICE_GPU)
        got = ps.output()
    finally:
        ps.close()
    np.testing.assert_array_equal(got, np.cumsum(x.astype(np.int64)).astype(np.int32))


def test_prefix_sum_both_devices(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(0)
    n = 2_000_000
    x = np.random.default_rng(4).integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32)
    ps = U.PrefixSum(n, 100_000)
    try:
        ps.set_input(x)
        ps.run(U.DEVICE_BOTH)
        got = ps.output()
    finally:
        ps.close()
    np.testing.assert_array_equal(got, O.prefix_sum(x))
