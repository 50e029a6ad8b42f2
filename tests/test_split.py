"""Subtask splitting on the CPU stubs (no GPU needed).

A task submitted with canSplitCpuSubtasks runs each subtask as pieces on the CPU stubs, every piece
called with pmSubtaskInfo.splitInfo = {splitId, splitCount}, and the subtask completes with its
last piece (reference pmSubtaskSplitter.cpp:49-140: one split group of all CPU stubs, splitCount =
the group's size). The probe is tests/native/split_probe.cpp; its pieces take the apps' row split
(matrixMultiplyBlas.cpp:67-88: equal parts, the last takes the rest, pieces beyond the row count
do nothing). A time-out while pieces are still queued drops them and fails the task with
pmTimerFailure, without waiting for them. The imageFiltering and matrixMultiply apps split their
CPU subtasks the same way and stay bit-exact against the oracle.
"""
import ctypes as C
import time
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle_lib as O

PROBE = Path(__file__).resolve().parent / "native" / "libsplit_probe.so"
PM_TIMER_FAILURE = 10
CPU_STUBS = 4            # the session runtime's CPU stubs (conftest: UA_TEST_CPU_STUBS, default 4)


@pytest.fixture(scope="module")
def lib(runtime):
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.split_run.restype = C.c_int
    L.split_run.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                            C.c_uint32, C.POINTER(C.c_uint32)]
    yield L
    runtime.set_split()


def _run(lib, subtasks, rows, pieces, sleep_ms=0, timeout=-1):
    out = np.zeros(subtasks * rows, dtype=np.uint32)
    cap = 3 * subtasks * max(pieces, CPU_STUBS, 1) + 3
    log = np.zeros(cap, dtype=np.uint32)
    n = C.c_uint32(0)
    rc = lib.split_run(subtasks, rows, pieces, sleep_ms, timeout, out.ctypes.data, log.ctypes.data, cap, C.byref(n))
    assert n.value <= cap
    return rc, out, log[: n.value].reshape(-1, 3)


def _expected(subtasks, rows):
    s = np.repeat(np.arange(subtasks, dtype=np.uint32), rows)
    r = np.tile(np.arange(rows, dtype=np.uint32), subtasks)
    return (s << 16) | r


@pytest.mark.parametrize("pieces", [0, 1, 3, 7, 150])
def test_split_pieces_cover_every_row_once(lib, pieces):
    subtasks, rows = 5, 100
    rc, out, log = _run(lib, subtasks, rows, pieces)
    assert rc == 0
    np.testing.assert_array_equal(out, _expected(subtasks, rows))
    count = CPU_STUBS if pieces == 0 else pieces
    if count < 2:
        # no splitting: one call per subtask, splitCount 0
        assert sorted(log[:, 0].tolist()) == list(range(subtasks)) and np.all(log[:, 1:] == 0)
        return
    assert np.all(log[:, 2] == count)
    for s in range(subtasks):
        ids = sorted(log[log[:, 0] == s, 1].tolist())
        assert ids == list(range(count)), (s, ids)          # every piece once, none twice


def test_split_ragged_rows(lib):
    # 7 rows in 3 pieces: 2, 2, 3; 1 row in 4 pieces: only piece 0 has a row
    for rows, pieces in [(7, 3), (1, 4), (5, 5)]:
        rc, out, log = _run(lib, 3, rows, pieces)
        assert rc == 0
        np.testing.assert_array_equal(out, _expected(3, rows))
        assert log.shape[0] == 3 * pieces


def test_split_timeout_drops_queued_pieces(lib):
    # 6 subtasks x 8 pieces x 150 ms on 4 stubs take ~1.8 s: a 1 s time-out fires with pieces queued
    t0 = time.monotonic()
    rc, out, log = _run(lib, 6, 64, 8, sleep_ms=150, timeout=1)
    took = time.monotonic() - t0
    assert rc == PM_TIMER_FAILURE
    assert log.shape[0] < 6 * 8                  # some pieces never ran
    assert took < 1.0 + 2 * 0.15 + 1.0           # the wait ended with the pieces already running
    # no piece ran twice, and the runtime still runs tasks afterwards
    keys = set(map(tuple, log[:, :2].tolist()))
    assert len(keys) == log.shape[0]
    rc, out, _ = _run(lib, 2, 16, 4)
    assert rc == 0
    np.testing.assert_array_equal(out, _expected(2, 16))


@pytest.mark.parametrize("pieces", [1, 3, 1000])
def test_imagefilter_split_cpu_matches_oracle(runtime, pieces):
    U = runtime
    rng = np.random.default_rng(11)
    H, W, r = 70, 90, 2
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
    U.set_split(cpu=True, cpu_pieces=pieces)
    try:
        f = U.ImageFilter(W, H, w, r, mode=0, shift=4, tile=32)
        try:
            f.set_input(img)
            f.run(U.DEVICE_CPU)
            got = f.output()
        finally:
            f.close()
    finally:
        U.set_split()
    assert np.array_equal(got, O.imagefilter(img, w, r, 0, 4))


@pytest.mark.parametrize("pieces", [1, 5, 100])
def test_matmul_split_cpu_integer_exact(runtime, pieces):
    U = runtime
    rng = np.random.default_rng(12)
    n, bs = 128, 32
    A = rng.integers(-8, 9, (n, n)).astype(np.float32)
    B = rng.integers(-8, 9, (n, n)).astype(np.float32)
    U.set_split(cpu=True, cpu_pieces=pieces)
    try:
        mm = U.MatMul(n, bs, 0)
        try:
            mm.set_inputs(A, B)
            mm.run(U.DEVICE_CPU)
            Cm = mm.output()
        finally:
            mm.close()
    finally:
        U.set_split()
    assert np.array_equal(Cm, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))


def test_split_factors_reject_bad_counts(runtime):
    with pytest.raises(Exception):
        runtime.set_split(cpu_pieces=-1)
    runtime.set_split()
