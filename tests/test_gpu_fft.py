"""fft on the GPU: uh_fft_c2c_f32 (the cuFFT replacement) directly, and the app's row and column
tasks through the HIP stub in all three modes, against the FP64 oracle (oracle.fft2, pinned in
test_fft.py) and numpy's complex128 FFT for the 1-D sequences.

Tolerance (float outputs): ||X - X64||_F / ||X64||_F <= 1e-5 per transform; the measured error
is printed. Views: PLAIN mode gives the column task COMPACT views (R columns per packed row),
RESIDENT / BATCHED give NATURAL views of the HBM mirror (row stride = matrix row) unless COMPACT
is forced; the callback reads the stride from the view type.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_lib as O
from tests.test_fft import TOL, _ramp, _rand, _rel, ramp_closed_form

pytestmark = pytest.mark.gpu


def _fft_dev(x, seq_axis, n, count, in_place=False, pad=0):
    """Runs uh_fft_c2c_f32 over x (complex64, 2-D) along the axis that is not seq_axis."""
    from unicorn_amd import hip as H
    rows, cols = x.shape
    d_in = H.DeviceBuffer.from_array(x)
    d_out = d_in if in_place else H.DeviceBuffer(x.nbytes)
    if seq_axis == 0:           # sequences are rows: element stride 1
        args = (cols, 1)
    else:                       # sequences are columns: sequence stride 1, element stride = row
        args = (1, cols)
    try:
        H.check(H.lib().uh_fft_c2c_f32(d_in.ptr, args[0], args[1], d_out.ptr, args[0], args[1], n, count, None),
                "uh_fft_c2c_f32")
        H.check(H.lib().uh_stream_sync(None), "sync")
        return d_out.to_array(x.shape, np.complex64)
    finally:
        d_in.free()
        if not in_place:
            d_out.free()


@pytest.mark.parametrize("log2n", list(range(1, 14)))
def test_kernel_rows(gpu_runtime, log2n):
    n = 1 << log2n
    count = max(3, min(300, (1 << 20) // n)) + 1          # not a multiple of the rows per workgroup
    x = _rand(count, n, log2n)
    got = _fft_dev(x, 0, n, count)
    want = np.fft.fft(x.astype(np.complex128), axis=1)
    err = _rel(got, want)
    print(f"rows n={n} count={count} err={err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("log2n", list(range(1, 14)))
def test_kernel_columns(gpu_runtime, log2n):
    n = 1 << log2n
    cols = 37                                              # not a multiple of the columns per workgroup
    x = _rand(n, cols, 100 + log2n)
    got = _fft_dev(x, 1, n, cols)
    want = np.fft.fft(x.astype(np.complex128), axis=0)
    err = _rel(got, want)
    print(f"cols n={n} count={cols} err={err:.2e}")
    assert err <= TOL


def test_kernel_in_place_and_partial_count(gpu_runtime):
    x = _rand(64, 4096, 7)
    got = _fft_dev(x, 0, 4096, 50, in_place=True)
    want = np.fft.fft(x.astype(np.complex128), axis=1)
    assert _rel(got[:50], want[:50]) <= TOL
    assert np.array_equal(got[50:], x[50:])                  # rows past count are untouched


def test_kernel_rejects_bad_arguments(gpu_runtime):
    from unicorn_amd import hip as H
    d = H.DeviceBuffer(1 << 16)
    try:
        for n in (0, 1, 3, 96, 16384):
            assert H.lib().uh_fft_c2c_f32(d.ptr, n, 1, d.ptr, n, 1, n, 1, None) != 0
        assert H.lib().uh_fft_c2c_f32(d.ptr, 64, 1, d.ptr, 64, 2, 64, 1, None) != 0   # in place, strides differ
        assert H.lib().uh_fft_c2c_f32(d.ptr + 4, 64, 1, d.ptr + 4, 64, 1, 64, 1, None) != 0   # misaligned
    finally:
        d.free()


def _run(U, x, R, device):
    f = U.Fft2D(x.shape[0], x.shape[1], R)
    try:
        f.set_input(x)
        f.run(device)
        return f.output()
    finally:
        f.close()


@pytest.mark.parametrize("hip_mode,compact", [(0, False), (1, False), (2, False), (1, True)])
@pytest.mark.parametrize("shape,R", [((256, 256), 64), ((512, 2048), 256), ((2048, 128), 128)])
def test_app_gpu(gpu_runtime, hip_mode, compact, shape, R):
    U = gpu_runtime
    U.set_hip_mode(hip_mode, compact=compact)
    try:
        x = _rand(*shape, 11)
        err = _rel(_run(U, x, R, U.DEVICE_GPU), O.fft2(x))
    finally:
        U.set_hip_mode(hip_mode)
    print(f"app {shape} R={R} mode={hip_mode} compact={compact} err={err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("hip_mode", [0, 1])
def test_app_reference_default_size(gpu_runtime, hip_mode):
    """The reference's default run: 4096 x 4096, 512 rows per subtask, input (i, i + 1)."""
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    R = Cc = 4096
    got = _run(U, _ramp(R, Cc), 512, U.DEVICE_GPU)
    err = _rel(got, ramp_closed_form(R, Cc))
    print(f"reference input 4096^2 mode={hip_mode} err={err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("hip_mode", [0, 2])
def test_app_both_devices(gpu_runtime, hip_mode):
    """CPU and HIP stubs share the tasks (CPU subtasks split, multi-assign on)."""
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    U.set_split(cpu=True)
    U.set_multi_assign(True)
    try:
        x = _rand(1024, 1024, 13)
        err = _rel(_run(U, x, 64, U.DEVICE_BOTH), O.fft2(x))
    finally:
        U.set_split()
    print(f"both mode={hip_mode} err={err:.2e}")
    assert err <= TOL
