"""matrixTranspose through the HIP stub, bit-exact against the oracle.

Out of place in every stub mode (PLAIN stages COMPACT views; RESIDENT / BATCHED transpose
between HBM mirrors); in place, where the runtime holds the writes back until the task ends and
the GPU callback locates the read and write rows of each block in the COMPACT view of the one
READ_WRITE space. The last case is the reference app's default size (2^14 x 2^12,
matrixTranspose.h:36-37): 16 blocks of 2048 x 2048.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _run(U, m, inplace, device):
    rows, cols = m.shape
    t = U.MatTranspose(rows, cols, inplace)
    try:
        t.set_input(m)
        t.run(device)
        return t.output()
    finally:
        t.close()


def _mat(rows, cols, seed):
    return np.random.default_rng(seed).integers(-(1 << 31), 1 << 31, (rows, cols), dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("hip_mode,compact", [(0, False), (1, False), (1, True), (2, False)])
@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("rows,cols", [(256, 256), (512, 128), (192, 576), (100, 100)])
def test_transpose_gpu(gpu_runtime, hip_mode, compact, inplace, rows, cols):
    U = gpu_runtime
    U.set_hip_mode(hip_mode, compact=compact)
    m = _mat(rows, cols, rows + 7 * cols + hip_mode)
    np.testing.assert_array_equal(_run(U, m, inplace, U.DEVICE_GPU), O.transpose(m))


@pytest.mark.parametrize("inplace", [False, True])
def test_transpose_both_devices(gpu_runtime, inplace):
    U = gpu_runtime
    U.set_hip_mode(0)
    m = _mat(1024, 256, 3)
    np.testing.assert_array_equal(_run(U, m, inplace, U.DEVICE_BOTH), O.transpose(m))


def test_transpose_reference_default_size(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(2)
    m = _mat(1 << 14, 1 << 12, 11)
    got = _run(U, m, False, U.DEVICE_GPU)
    np.testing.assert_array_equal(got, m.T)
