"""matrixTranspose app on the CPU stubs (no GPU needed), bit-exact against the oracle.

Out of place: COMPACT views of the input and output (reference matrixTranspose.cpp:474-478).
In place: one READ_WRITE space; subtasks read the input layout and write the output layout of
the same space, so for a non-square matrix their reads overlap other subtasks' writes and the
runtime must hold the writes back until the task ends (the reference's shadow memory for
non-disjoint READ_WRITE spaces). Blocks are min(rows, cols, 2048) square (matrixTranspose.cpp:52-61).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("rows,cols", [(64, 64), (96, 32), (32, 160), (7, 7), (1, 5)])
def test_transpose_cpu(runtime, rows, cols, inplace):
    U = runtime
    m = np.random.default_rng(rows * 31 + cols).integers(-(1 << 31), 1 << 31, (rows, cols), dtype=np.int64)
    m = m.astype(np.int32)
    t = U.MatTranspose(rows, cols, inplace)
    try:
        t.set_input(m)
        t.run(U.DEVICE_CPU)
        got = t.output()
    finally:
        t.close()
    np.testing.assert_array_equal(got, O.transpose(m))


def test_transpose_rejects_ragged_blocks(runtime):
    with pytest.raises(Exception):
        runtime.MatTranspose(30, 45)          # 45 is not a multiple of min(30, 45)
