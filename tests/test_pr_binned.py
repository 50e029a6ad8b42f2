"""The binned (propagation-blocked) pageRank layout and its CPU twin (host code, no GPU).

The layout is pinned against the out-link CSR it is built from: reading every run back (source =
its chunk's first page + src16, slot = run_pos + k, destination = its bin's first page + dst16)
must give exactly the CSR's links, once each; bins hold only links into their unit, padded to a
multiple of 4 with destination 0; units are subtask-aligned and at most 16384 pages.

The CPU twin (the app's CPU callback over the binned layout) adds fixed-point values, so its ranks
do not depend on summation order: they must match the FP64 recurrence to float rounding and the
oracle (fp32, push order) to 1e-5 relative L1, and a planted wrong damping or a dropped link must
be caught.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

UNIT_MAX = 16384


@pytest.fixture(scope="module")
def U():
    import unicorn_amd as U
    return U


def _links_back(L):
    """(src, dst, slot) of every link, read back from the runs."""
    us = L["unit_start"].astype(np.int64)
    bs = L["bin_start"].astype(np.int64)
    srcs, dsts, slots = [], [], []
    for c in range(L["units"]):
        for r in range(int(L["run_ptr"][c]), int(L["run_ptr"][c + 1])):
            o0, o1, p = int(L["run_off"][r]), int(L["run_off"][r + 1]), int(L["run_pos"][r])
            assert o1 > o0                                       # runs are never empty
            b = int(np.searchsorted(bs, p, side="right") - 1)
            assert p + (o1 - o0) <= bs[b + 1]                   # a run stays inside its bin
            slot = np.arange(p, p + (o1 - o0))
            srcs.append(us[c] + L["src16"][o0:o1].astype(np.int64))
            dsts.append(us[b] + L["dst16"][slot].astype(np.int64))
            slots.append(slot)
    cat = lambda v: np.concatenate(v) if v else np.zeros(0, np.int64)
    return cat(srcs), cat(dsts), cat(slots)


def _check(U, rp, col, per, cus=256):
    n = rp.size - 1
    L = U.pr_binned_layout(rp, col, per, cus)
    assert L is not None
    us = L["unit_start"].astype(np.int64)
    assert us[0] == 0 and us[-1] == n and np.all(np.diff(us) >= 1) and np.all(np.diff(us) <= UNIT_MAX)
    # subtask boundaries are unit boundaries
    assert set(range(0, n, per)) <= set(us[:-1].tolist())
    bs = L["bin_start"].astype(np.int64)
    assert bs[0] == 0 and np.all(bs % 4 == 0) and bs[-1] == L["slots"]
    indeg = np.bincount(col.astype(np.int64), minlength=n)
    per_bin = np.add.reduceat(indeg, us[:-1]) if n else np.zeros(0)
    np.testing.assert_array_equal(np.diff(bs), (per_bin + 3) // 4 * 4)
    np.testing.assert_array_equal(L["deg"], np.diff(rp.astype(np.int64)))
    src, dst, slot = _links_back(L)
    assert slot.size == col.size and np.unique(slot).size == slot.size     # each slot once
    want_src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    got = np.lexsort((dst, src))
    ref = np.lexsort((col.astype(np.int64), want_src))
    np.testing.assert_array_equal(src[got], want_src[ref])
    np.testing.assert_array_equal(dst[got], col.astype(np.int64)[ref])
    # padding: every slot no run fills has destination 0
    pad = np.ones(L["slots"], dtype=bool)
    pad[slot] = False
    assert np.all(L["dst16"][pad] == 0)
    return L


def test_binned_layout_webgen(U):
    rp, col = U.webgen(60_000, 9)
    L = _check(U, rp, col, 25_000)
    assert L["subtasks"] == 3


def test_binned_layout_tiny_and_ragged(U):
    rp = np.array([0, 1, 2, 2], dtype=np.uint32)            # 3 pages, one dangling
    col = np.array([1, 0], dtype=np.uint32)
    _check(U, rp, col, 2)
    rp, col = U.webgen(40_000, 4)
    _check(U, rp, col, 40_000, cus=8)                         # one subtask, units capped by run length
    _check(U, rp, col, 7_777, cus=0)                          # no rounding to whole rounds


def test_binned_units_fill_whole_rounds_when_runs_stay_long(U):
    rp, col = U.webgen(300_000, 5)
    L = U.pr_binned_layout(rp, col, 100_000, cus=16)
    assert L["units"] % 16 == 0
    assert L["runs"] / L["units"] ** 2 > 0.5                   # most (chunk, bin) pairs hold links


def _rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def test_binned_cpu_matches_oracle_and_fp64(runtime):
    U = runtime
    rp, col = U.webgen(50_000, 21)
    U.pr_set_layout("binned")
    try:
        pr = U.PageRank(rp, col, 12_000)
        assert pr.layout == "binned"
        pr.run(12, U.DEVICE_CPU)
        y = pr.ranks()
        pr.close()
    finally:
        U.pr_set_layout("binned")
    y64 = O.pagerank_f64(rp, col, 12)
    yo = O.pagerank(rp, col, 12, 12_000)
    # the error left is the per-iteration float rounding of y and w, which the reference's own
    # algorithm carries too: no worse than twice the oracle's distance to FP64
    assert _rel_l1(y, y64) <= max(2 * _rel_l1(yo, y64), 1e-7) and _rel_l1(y, y64) <= 1e-6
    assert _rel_l1(y, yo) <= 1e-5
    # planted mistakes the comparison catches: wrong damping, a dropped link
    assert _rel_l1(y, O.pagerank_f64(rp, col, 12) * (0.84 / 0.85) ** 12) > 1e-3
    col2 = col.copy()
    col2[len(col2) // 2] = (col2[len(col2) // 2] + 1) % (rp.size - 1)
    assert _rel_l1(y, O.pagerank_f64(rp, col2, 12)) > 1e-7


def test_binned_cpu_ring_and_star(runtime):
    U = runtime
    n = 1000
    rp = np.arange(n + 1, dtype=np.uint32)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)          # ring
    U.pr_set_layout("binned")
    try:
        pr = U.PageRank(rp, col, 300)
        pr.run(3, U.DEVICE_CPU)
        y = pr.ranks()
        pr.close()
        assert np.all(y == np.float32(0.85) ** 3) or np.allclose(y, 0.85 ** 3, rtol=1e-6)
        rp = np.concatenate([[0, 0], np.arange(1, n)]).astype(np.uint32)   # star: pages 1.. -> 0; hub dangling
        col = np.zeros(n - 1, dtype=np.uint32)
        pr = U.PageRank(rp, col, 256)
        pr.run(1, U.DEVICE_CPU)
        y = pr.ranks()
        pr.close()
    finally:
        U.pr_set_layout("binned")
    assert abs(float(y[0]) - 0.85 * (n - 1)) <= 1e-6 * 0.85 * (n - 1)
    assert np.all(y[1:] == 0)
