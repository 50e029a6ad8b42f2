"""Out-of-memory back-off of the HIP stub's per-subtask path (reference
pmExecutionStub.cpp:2403-2415): a device allocation that does not fit the HBM budget waits for the
stub's oldest in-flight subtask and retries, instead of failing the task.

With a budget that holds one subtask's staging buffers but not two, a PLAIN imageFiltering task of
16 subtasks (up to 4 in flight) completes bit-exact against the oracle and the stub reports retried
allocations. A subtask whose buffers alone exceed the budget still fails the task with
pmMemoryError (nothing in flight to wait for).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _run(U, budget, seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, (2048, 2048, 3), dtype=np.uint8)
    w = rng.integers(-4, 5, 25).astype(np.int8)
    U.set_hip_mode(0)
    U.set_hbm_budget(budget)
    f = U.ImageFilter(2048, 2048, w, 2, mode=1, shift=4, tile=512)
    try:
        f.set_input(img)
        U.reset_stats()
        f.run(U.DEVICE_GPU)
        st = U.stats()
        got = f.output()
    finally:
        f.close()
        U.set_hbm_budget(0)
    return got, st, O.imagefilter(img, w, 2, 1, 4)


def test_oom_backoff_waits_for_oldest_subtask(gpu_runtime):
    got, st, want = _run(gpu_runtime, 5 << 19, 3)     # 2.5 MiB: one subtask (1.6 MB) fits, two do not
    assert np.array_equal(got, want)
    assert st["subtasks"] == 16
    assert st["oom_waits"] >= 1, st


def test_oom_without_inflight_subtasks_fails(gpu_runtime):
    with pytest.raises(RuntimeError):
        _run(gpu_runtime, 1 << 20, 4)


def test_no_waits_with_room(gpu_runtime):
    got, st, want = _run(gpu_runtime, 0, 5)
    assert np.array_equal(got, want)
    assert st["oom_waits"] == 0
