"""The PLAIN stub's read-only device cache (reference pmCudaCacheKey, pmDataTypes.h:295-387, an
LRU of device copies keyed by subscription; purged when the address space is written,
pmAddressSpace.cpp:591-594). Entries are keyed by (address space, write version, records): a
repeated read of the same records hits, within a task and across tasks, and any write to the
space (here through pmGetRawMemPtr) makes the old entries unreachable, so the next task reads the
new data.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_readonly_cache_hits_and_write_invalidation(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(0)
    n, bs = 512, 128                     # 16 subtasks; each A row panel and B column panel is read 4 times
    rng = np.random.default_rng(8)
    A = rng.integers(-8, 9, (n, n)).astype(np.float32)
    B = rng.integers(-8, 9, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    try:
        mm.set_inputs(A, B)
        U.reset_stats()
        mm.run(U.DEVICE_GPU)
        first = U.stats()["cache_hits"]
        assert np.array_equal(mm.output(), (A.astype(np.float64) @ B).astype(np.float32))
        U.reset_stats()
        mm.run(U.DEVICE_GPU)
        second = U.stats()["cache_hits"]
        # within the first task: 3 of the 4 reads of each of 4 A panels and 4 B panels hit
        assert first == 24, first
        # the second task finds all 32 panel reads cached
        assert second == 32, second
        A2, B2 = -A, B[::-1].copy()
        mm.set_inputs(A2, B2)            # host writes: both spaces get a new version
        U.reset_stats()
        mm.run(U.DEVICE_GPU)
        assert U.stats()["cache_hits"] == 24
        assert np.array_equal(mm.output(), (A2.astype(np.float64) @ B2).astype(np.float32))
    finally:
        mm.close()
