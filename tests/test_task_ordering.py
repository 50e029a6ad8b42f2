"""Ordering of tasks that share an address space (CPU stubs only, no GPU).

A writer and a reader of the same space submitted back to back without waiting: the reader must
see the writer's data and start after the writer's end, as the reference orders tasks by locking
their address spaces (pmAddressSpace.cpp:500-530, pmTask.cpp:176-193). A reader submitted with
pmSubmitTaskWithOptions(unordered) starts at once and sees the old data. Tasks that only read a
shared space run at the same time. The probe is tests/native/ordering_probe.cpp.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

PROBE = Path(__file__).resolve().parent / "native" / "libordering_probe.so"


@pytest.fixture(scope="module")
def lib(runtime):
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.ord_writer_reader.restype = C.c_int
    L.ord_writer_reader.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.POINTER(C.c_double)]
    L.ord_readers_overlap.restype = C.c_int
    L.ord_readers_overlap.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]
    return L


def _writer_reader(lib, unordered, n=5000, subtasks=3, sleep_ms=150):
    out = np.zeros(n * subtasks, dtype=np.uint32)
    gap = C.c_double(0)
    assert lib.ord_writer_reader(unordered, n, subtasks, 0xC0FFEE, sleep_ms, out.ctypes.data, C.byref(gap)) == 0
    return out, gap.value


def test_dependent_reader_waits_for_writer(lib):
    out, gap = _writer_reader(lib, 0)
    assert np.all(out == 0xC0FFEE)
    assert gap >= 0.0


def test_unordered_reader_does_not_wait(lib):
    out, gap = _writer_reader(lib, 1)
    assert gap < 0.0                       # started while the writer slept
    assert np.all(out == 0)                # and copied the data from before the write


def test_readers_of_a_shared_space_run_together(lib, runtime):
    m = C.c_int(0)
    assert lib.ord_readers_overlap(3, 200, C.byref(m)) == 0
    assert m.value >= 2
