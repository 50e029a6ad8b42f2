"""Multi-assign on one host: a task that both stub kinds can run (DEVICE_BOTH) and that allows
multi-assign (pmTaskDetails::multiAssignEnabled, the reference's default) lets the idle HIP stub
run a second copy of a subtask that a CPU stub is still executing (reference
pmExecutionStub.cpp:1058-1170, at most one extra copy). The copy that finishes first commits
its output and completes the subtask; the other's output is dropped.

The probe (tests/native/ma_probe.cpp) makes the CPU callback slow and the GPU callback fast, so
the HIP stub finishes the other subtasks and then copies the ones the CPU stubs hold; its copies
win, and the output must be exact. imageFiltering with DEVICE_BOTH must stay bit-exact against
the oracle with copies in play; with multi-assign off no copy may run.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu
PROBE = Path(__file__).resolve().parent / "native" / "libma_probe.so"


@pytest.fixture
def both(gpu_runtime):
    U = gpu_runtime
    U.set_split(cpu=False)          # split subtasks are not multi-assigned
    yield U
    U.set_split()
    U.set_multi_assign(True)


@pytest.fixture(scope="module")
def probe(gpu_runtime):
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.ma_run.restype = C.c_int
    L.ma_run.argtypes = [C.c_uint32] * 5 + [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    return L


def _probe(L, subtasks, n, sleep_ms, base, ma):
    out = np.zeros(subtasks * n, np.uint32)
    cp, w = C.c_uint64(), C.c_uint64()
    rc = L.ma_run(subtasks, n, sleep_ms, base, ma, out.ctypes.data, C.byref(cp), C.byref(w))
    assert rc == 0
    np.testing.assert_array_equal(out, np.repeat(np.arange(subtasks, dtype=np.uint32) * 7 + 1 + base, n))
    return cp.value, w.value


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_gpu_copies_slow_cpu_subtasks(gpu_runtime, probe, hip_mode):
    gpu_runtime.set_hip_mode(hip_mode)
    copies, wins = _probe(probe, 40, 5000, 150, hip_mode, 1)
    assert copies >= 1 and 1 <= wins <= copies, (copies, wins)


def test_no_copies_without_multi_assign_probe(gpu_runtime, probe):
    gpu_runtime.set_hip_mode(0)
    assert _probe(probe, 12, 1000, 20, 5, 0) == (0, 0)


def _run_if(U, H, W, r, tile, seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
    f = U.ImageFilter(W, H, w, r, mode=0, shift=4, tile=tile)
    try:
        f.set_input(img)
        U.reset_stats()
        f.run(U.DEVICE_BOTH)
        st = U.stats()
        got = f.output()
    finally:
        f.close()
    assert np.array_equal(got, O.imagefilter(img, w, r, 0, 4))
    return st


@pytest.mark.parametrize("hip_mode", [0, 2])
def test_imagefilter_both_with_multi_assign(both, hip_mode):
    U = both
    U.set_hip_mode(hip_mode)
    U.set_multi_assign(True)
    st = _run_if(U, 2048, 2048, 2, 256, 5 + hip_mode)
    assert st["ma_wins"] <= st["ma_copies"]


def test_no_copies_without_multi_assign(both):
    U = both
    U.set_hip_mode(0)
    U.set_multi_assign(False)
    st = _run_if(U, 1024, 1024, 2, 256, 9)
    assert st["ma_copies"] == 0, st


def test_multi_assign_repeated_tasks_release_cleanly(both):
    """Tasks released while losing CPU copies still run: release waits for them, and the next
    task on fresh address spaces is unaffected."""
    U = both
    U.set_hip_mode(0)
    U.set_multi_assign(True)
    for k in range(3):
        _run_if(U, 1024, 1536, 1, 256, 20 + k)
