"""Built-in reductions and data redistribution through the HIP stub.

The probe task of tests/test_reductions.py runs with its kernel-kind GPU callback: in PLAIN mode
the subtasks' outputs come back to host shadows and fold there; in RESIDENT and BATCHED mode the
shadows stay in HBM and fold with uh_reduce, for all ten operations over all six element types,
equal to numpy's fold. Redistribution: the GPU callback writes its chunk lengths into a
SUBTASK_TO_POST_SUBTASK scratch buffer that the runtime copies back for the dataRedistribution
callback; the redistributed space must equal the concatenation by (order, subtask).
"""
import numpy as np
import pytest

from test_reductions import TYPES, expected_redistribution, fold, gen, load, run_reduce, run_redistribute

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_runtime):
    return load(gpu_runtime)


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("ti", range(6))
def test_builtin_reductions_on_gpu(gpu_runtime, lib, hip_mode, ti):
    gpu_runtime.set_hip_mode(hip_mode)
    n, subtasks = 3000, 6
    for op in range(10):
        rc, got = run_reduce(lib, op, ti, subtasks, n, 1)
        assert rc == 0, (op, ti)
        want = fold(op, TYPES[ti], [gen(op, TYPES[ti], s, n) for s in range(subtasks)])
        np.testing.assert_array_equal(got, want, err_msg=f"op {op} type {ti} mode {hip_mode}")


@pytest.mark.parametrize("device", [1, 2])
@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_redistribution_on_gpu(gpu_runtime, lib, device, hip_mode):
    """PLAIN stages the subtasks through host memory and redistributes there; a GPU-only task in
    the resident modes keeps its output in HBM and is redistributed there (one copy kernel)."""
    gpu_runtime.set_hip_mode(hip_mode)
    subtasks, n, orders = 8, 1000, 5
    rc, out, meta = run_redistribute(lib, subtasks, n, orders, device)
    assert rc == 0
    body, want_meta = expected_redistribution(subtasks, n, orders)
    np.testing.assert_array_equal(out[:body.size], body)
    assert np.all(out[body.size:] == 0)
    assert meta == want_meta


@pytest.fixture
def sentinel_all(gpu_runtime):
    """Sentinel compression for outputs of any size (restored afterwards)."""
    prev = gpu_runtime.set_sentinel_threshold(0)
    yield gpu_runtime
    gpu_runtime.set_sentinel_threshold(prev)


@pytest.mark.parametrize("ti", range(6))
def test_sentinel_compressed_reduction_on_gpu(sentinel_all, lib, ti):
    """PLAIN mode with the threshold lowered to 0: each subtask writes its task's only address
    space whole with about one element in 16 non-zero, so its output goes back sentinel-compressed
    (reference pmExecutionStub.cpp:3850-3950) and is expanded into the host shadow before the fold.
    The fold equals numpy's, and far fewer bytes cross PCIe than the outputs hold."""
    U = sentinel_all
    U.set_hip_mode(0)
    n, subtasks = 20000, 6
    t = TYPES[ti]
    for op in range(10):
        U.reset_stats()
        out = np.zeros(n, dtype=t)
        rc = lib.rp_reduce_ex(op, ti, subtasks, n, 1, 1, out.ctypes.data)
        st = U.stats()
        assert rc == 0, (op, ti)
        want = fold(op, t, [gen(op, t, s, n, sparse=True) for s in range(subtasks)])
        np.testing.assert_array_equal(out, want, err_msg=f"op {op} type {ti}")
        assert st["sentinel_compressed"] == subtasks, st
        assert st["d2h_bytes"] < subtasks * n * np.dtype(t).itemsize // 2, st


def test_sentinel_dense_output_and_default_threshold(sentinel_all, lib):
    """A dense output (6 in 7 elements non-zero) is sent whole; at the reference's 1 GiB threshold
    a small sparse output is sent whole too."""
    U = sentinel_all
    U.set_hip_mode(0)
    n, subtasks = 20000, 6
    U.reset_stats()
    rc, got = run_reduce(lib, 0, 0, subtasks, n, 1)
    assert rc == 0
    np.testing.assert_array_equal(got, fold(0, TYPES[0], [gen(0, TYPES[0], s, n) for s in range(subtasks)]))
    assert U.stats()["sentinel_compressed"] == 0
    U.set_sentinel_threshold(1 << 30)
    U.reset_stats()
    out = np.zeros(n, dtype=TYPES[0])
    assert lib.rp_reduce_ex(0, 0, subtasks, n, 1, 1, out.ctypes.data) == 0
    np.testing.assert_array_equal(out, fold(0, TYPES[0], [gen(0, TYPES[0], s, n, sparse=True) for s in range(subtasks)]))
    assert U.stats()["sentinel_compressed"] == 0


def test_sentinel_kernel_matches_numpy(gpu_runtime):
    """uh_sentinel_compress against numpy: the count, the non-zero elements in index order and
    their indices, for every element type, at sizes around the 4096-element blocks; float -0.0
    compares equal to zero and is dropped, as in the reference; over the bound only the count."""
    import ctypes as C
    from unicorn_amd import hip as H
    rng = np.random.default_rng(7)
    types = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
    for ti, dt in enumerate(types):
        es = np.dtype(dt).itemsize
        for n in (1, 4095, 4096, 4097, 100003):
            a = np.zeros(n, dtype=dt)
            mask = rng.random(n) < 0.1
            a[mask] = rng.integers(1, 1000, int(mask.sum())).astype(dt)
            if dt in (np.float32, np.float64) and n > 10:
                a[3] = -0.0
                a[5] = -2.5
            want_idx = np.nonzero(a != 0)[0].astype(np.uint32)
            d_in = H.DeviceBuffer.from_array(a)
            d_out = H.DeviceBuffer(8 + n * (es + 4))
            nz = C.c_uint64(0)
            try:
                H.check(H.lib().uh_sentinel_compress(d_in.ptr, n, ti, d_out.ptr, n, C.byref(nz), None), "compress")
                H.check(H.lib().uh_stream_sync(None), "sync")
                k = int(nz.value)
                assert k == want_idx.size, (ti, n)
                buf = d_out.to_array((8 + k * (es + 4),), np.uint8)
                assert int(buf[:8].view(np.uint64)[0]) == k
                np.testing.assert_array_equal(buf[8:8 + k * es].view(dt), a[want_idx])
                np.testing.assert_array_equal(buf[8 + k * es:].view(np.uint32), want_idx)
                if k:   # over the bound: the count only
                    nz2 = C.c_uint64(0)
                    H.check(H.lib().uh_sentinel_compress(d_in.ptr, n, ti, d_out.ptr, k - 1, C.byref(nz2), None), "count")
                    assert int(nz2.value) == k
            finally:
                d_in.free()
                d_out.free()
