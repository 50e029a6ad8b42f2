"""Multi-assign on one host: a task that both stub kinds can run (DEVICE_BOTH) and that allows
multi-assign (pmTaskDetails::multiAssignEnabled, the reference's default) lets the idle HIP stub
run a second copy of a subtask that a CPU stub is still executing (reference
pmExecutionStub.cpp:1058-1170, at most one extra copy). The copy that finishes first commits
its output and completes the subtask; the other's output is dropped. The task's output must
equal the oracle byte for byte either way, and with multi-assign off no copy may run.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def both(gpu_runtime):
    U = gpu_runtime
    U.set_split(cpu=False)          # split subtasks are not multi-assigned
    yield U
    U.set_split()
    U.set_multi_assign(True)


def _run(U, H, W, r, tile, seed):
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
def test_idle_gpu_copies_cpu_subtasks(both, hip_mode):
    U = both
    U.set_hip_mode(hip_mode)
    U.set_multi_assign(True)
    # 64 subtasks of 256 x 256: the CPU stubs each hold one for tens of milliseconds while the
    # HIP stub finishes the rest in about a millisecond, then copies what the CPU stubs still run
    st = _run(U, 2048, 2048, 2, 256, 5 + hip_mode)
    assert st["ma_copies"] >= 1, st
    assert st["ma_wins"] >= 1, st
    assert st["ma_wins"] <= st["ma_copies"]


def test_no_copies_without_multi_assign(both):
    U = both
    U.set_hip_mode(0)
    U.set_multi_assign(False)
    st = _run(U, 1024, 1024, 2, 256, 9)
    assert st["ma_copies"] == 0, st


def test_multi_assign_repeated_tasks_release_cleanly(both):
    """Tasks released while losing CPU copies still run: release waits for them, and the next
    task on fresh address spaces is unaffected."""
    U = both
    U.set_hip_mode(0)
    U.set_multi_assign(True)
    for k in range(3):
        _run(U, 1024, 1536, 1, 256, 20 + k)
