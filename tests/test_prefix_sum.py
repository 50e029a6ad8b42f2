"""prefixSum on the CPU stubs (no GPU needed), bit-exact against the oracle (serialPrefixSum over
int, wrap-around): the chunk-scan task, the host scan of the chunk totals and the add task
(reference prefixSum.cpp:225-290)."""
import numpy as np
import pytest

from oracle import oracle_lib as O


@pytest.mark.parametrize("n,chunk", [(10000, 1000), (10007, 1000), (999, 1000), (1, 5), (65536, 4096)])
def test_prefix_sum_cpu(runtime, n, chunk):
    U = runtime
    x = np.random.default_rng(n + chunk).integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32)
    ps = U.PrefixSum(n, chunk)
    try:
        ps.set_input(x)
        ps.run(U.DEVICE_CPU)
        got = ps.output()
    finally:
        ps.close()
    np.testing.assert_array_equal(got, O.prefix_sum(x))
