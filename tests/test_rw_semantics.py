"""READ_WRITE address spaces whose subtasks read what other subtasks write (CPU stubs only).

The probe (tests/native/rw_probe.cpp) runs a task in which subtask s reads block s+1 (cyclic)
and writes block s. Unless the task declares its reads and writes disjoint across subtasks
(pmTaskMem::disjointReadWritesAcrossSubtasks), every subtask must read the task's input, so the
result is exactly the input rotated by one block, plus one, whatever order the stubs run the
subtasks in: the writes are held back until the task ends (reference
pmTask::DoSubtasksNeedShadowMemory, pmTask.cpp:749-754). Both NATURAL and COMPACT views.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

PROBE = Path(__file__).resolve().parent / "native" / "librw_probe.so"


@pytest.fixture(scope="module")
def lib(runtime):
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.rw_shift.restype = C.c_int
    L.rw_shift.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_void_p]
    return L


@pytest.mark.parametrize("compact", [0, 1])
@pytest.mark.parametrize("n,m", [(16, 1000), (5, 1), (2, 4096)])
def test_overlapping_reads_see_task_input(lib, compact, n, m):
    rng = np.random.default_rng(n * 31 + m + compact)
    data = rng.integers(0, 1 << 30, n * m, dtype=np.uint32)
    want = (np.roll(data.reshape(n, m), -1, axis=0) + 1).reshape(-1)
    for _ in range(3):                       # stub interleavings differ between runs
        got = data.copy()
        assert lib.rw_shift(n, m, 0, compact, got.ctypes.data) == 0
        np.testing.assert_array_equal(got, want)


def test_single_subtask_reads_its_own_writes_target(lib):
    # one subtask reading and writing the same block: the write lands at the end
    data = np.arange(64, dtype=np.uint32)
    got = data.copy()
    assert lib.rw_shift(1, 64, 0, 0, got.ctypes.data) == 0
    np.testing.assert_array_equal(got, data + 1)
