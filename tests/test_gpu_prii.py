"""pageRankInvertedIndex on the GPU: uh_invert_index (bit-exact against the oracle's
invertIndex_cpu restatement, one subtask window at a time) and uh_prii_rank, then the app's GPU
and CPU+GPU runs in the three HIP stub modes. The reference's GPU callbacks for this app are empty
stubs (pageRankInvertedIndexCUDA.cu:92-100); these do the CPU callbacks' work.

Bars: the inverted index (integer data) is bit-exact. Ranks use float atomics on the GPU, so their
summation order is free: relative L1 against the oracle (the CPU callbacks' order) <= 1e-5, the
float-output bound; the measured error is printed.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O
from tests.test_prii import _dump, _graph

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _invert_dev(rec, first, count, total, P, groups):
    from unicorn_amd import hip as H
    max_out = rec.shape[1] - 1
    sub = np.ascontiguousarray(rec[first:first + count])
    d_dump = H.DeviceBuffer.from_array(sub)
    d_out = H.DeviceBuffer(max(1, count * max_out * 12))
    d_tot = H.DeviceBuffer(groups * 4)
    try:
        rc = H.lib().uh_invert_index(d_dump.ptr, count, max_out, first, total, P, groups, d_out.ptr, d_tot.ptr, None)
        if rc:
            return rc, None, None
        tot = d_tot.to_array(groups, np.uint32)
        n = int(tot.sum())
        e = d_out.to_array((n, 3), np.uint32) if n else np.zeros((0, 3), np.uint32)
        return 0, e, tot
    finally:
        d_dump.free()
        d_out.free()
        d_tot.free()


def _oracle_window(rec, s, P):
    """Entries of source subtask s in its own window order: its part of every group, groups ascending."""
    e, counts = O.prii_invert(rec, P)
    parts, start = [], 0
    for g, c in enumerate(counts):
        blk = e[start:start + c]
        parts.append(blk[blk[:, 0] // P == s])
        start += c
    return np.concatenate(parts) if parts else np.zeros((0, 3), np.uint32), np.array([len(p) for p in parts], np.uint32)


@pytest.mark.parametrize("pages,P,max_out", [(5000, 1000, 20), (300000, 100000, 50), (20000, 64, 8), (777, 777, 3)])
def test_invert_kernel_matches_oracle(gpu_runtime, pages, P, max_out):
    from unicorn_amd import hip as H
    row_ptr, col = _graph(pages, 5, mean_out=min(6.0, max_out / 2), max_out=max_out)
    rec = O.prii_records(row_ptr, col, max_out)
    groups = (pages + P - 1) // P
    for s in sorted({0, groups // 2, groups - 1}):
        first, count = s * P, min(P, pages - s * P)
        rc, e, tot = _invert_dev(rec, first, count, pages, P, groups)
        assert rc == 0, H.lib().uh_last_error()
        want, wtot = _oracle_window(rec, s, P)
        np.testing.assert_array_equal(tot, wtot)
        np.testing.assert_array_equal(e, want)


def test_invert_kernel_rejects_malformed(gpu_runtime):
    from unicorn_amd import hip as H
    row_ptr, col = _graph(1000, 6, max_out=10)
    rec = O.prii_records(row_ptr, col, 10)
    rec[5, 0] = 11                                               # more links than max_out
    rc, _, _ = _invert_dev(rec, 0, 1000, 1000, 500, 2)
    assert rc != 0 and b"malformed" in H.lib().uh_last_error()
    rec = O.prii_records(row_ptr, col, 10)
    p = int(np.nonzero(rec[:, 0])[0][0])
    rec[p, 1] = 1001                                             # a target beyond the last page
    rc, _, _ = _invert_dev(rec, 0, 1000, 1000, 500, 2)
    assert rc != 0


def test_rank_kernel(gpu_runtime):
    from unicorn_amd import hip as H
    pages, P = 50000, 50000
    row_ptr, col = _graph(pages, 8)
    rec = O.prii_records(row_ptr, col, 20)
    e, counts = O.prii_invert(rec, P)
    x = np.random.default_rng(1).random(pages).astype(np.float32)
    d_e, d_x, d_y = H.DeviceBuffer.from_array(e), H.DeviceBuffer.from_array(x), H.DeviceBuffer(pages * 4)
    try:
        H.check(H.lib().uh_prii_rank(d_e.ptr, e.shape[0], d_x.ptr, 1.0, 0.85, 0, d_y.ptr, pages, pages, None), "rank")
        H.check(H.lib().uh_stream_sync(None), "sync")
        y = d_y.to_array(pages, np.float32).astype(np.float64)
    finally:
        d_e.free(); d_x.free(); d_y.free()
    ref = np.zeros(pages)
    np.add.at(ref, e[:, 1], 0.85 * x[e[:, 0]].astype(np.float64) / e[:, 2])
    err = np.abs(y - ref).sum() / np.abs(ref).sum()
    print(f"rank kernel rel L1 {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("pages,per_file,P", [(20000, 1000, 5000), (3001, 250, 750)])
def test_app_gpu(gpu_runtime, tmp_path, hip_mode, pages, per_file, P):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    d, row_ptr, col, rec = _dump(tmp_path, pages, 12, per_file)
    a = U.PageRankInvertedIndex(str(d), P)
    try:
        a.run(U.DEVICE_GPU, 2)
        e, counts = a.index()
        ranks = a.ranks()
    finally:
        a.close()
    eo, co = O.prii_invert(rec, P)
    np.testing.assert_array_equal(counts, co)
    np.testing.assert_array_equal(e, eo)
    ro = O.prii_pagerank(eo, co, pages, P, 2).astype(np.float64)
    err = np.abs(ranks - ro).sum() / np.abs(ro).sum()
    print(f"app gpu mode={hip_mode} pages={pages} rel L1 {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("hip_mode", [0, 2])
def test_app_both_devices(gpu_runtime, tmp_path, hip_mode):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    U.set_multi_assign(True)
    pages, P = 40000, 2500
    d, row_ptr, col, rec = _dump(tmp_path, pages, 13, 500)
    a = U.PageRankInvertedIndex(str(d), P)
    try:
        a.run(U.DEVICE_BOTH, 3)
        e, counts = a.index()
        ranks = a.ranks()
    finally:
        a.close()
    eo, co = O.prii_invert(rec, P)
    np.testing.assert_array_equal(e, eo)
    ro = O.prii_pagerank(eo, co, pages, P, 3).astype(np.float64)
    err = np.abs(ranks - ro).sum() / np.abs(ro).sum()
    print(f"app both mode={hip_mode} rel L1 {err:.2e}")
    assert err <= TOL
