"""pageRank in the reference's decomposition on the CPU stubs: source-page subtasks push into
private full-length vectors and the built-in REDUCE_ADD folds them (pmReducer.cpp:593-605).

The oracle restates the same push + REDUCE_ADD; summation orders differ (the fold pairs subtasks
in completion order), hence relative L1 <= 1e-5. A ring pins the recurrence exactly (0.85^k per
page), and a dump loaded from disk must give the same ranks as the CSR it was written from.
"""
import numpy as np

from oracle import oracle_lib as O


def _rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def test_push_cpu_matches_oracle(runtime):
    U = runtime
    rp, col = U.webgen(40_000, 6)
    pr = U.PageRankPush(rp, col, 9_000)          # 5 subtasks, ragged last one
    try:
        pr.run(5, U.DEVICE_CPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 5, 9_000)) <= 1e-5


def test_push_cpu_ring_closed_form(runtime):
    U = runtime
    n = 3000
    rp = np.arange(n + 1, dtype=np.uint32)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)
    pr = U.PageRankPush(rp, col, 700)
    try:
        pr.run(3, U.DEVICE_CPU)
        np.testing.assert_allclose(pr.ranks(), np.float32(0.85) ** 3, rtol=1e-6)
    finally:
        pr.close()


def test_push_from_webdump_cpu(runtime, tmp_path):
    U = runtime
    rp, col = U.webgen(25_000, 9)
    U.webdump_write(tmp_path, rp, col, 100, 6000)
    a = U.PageRankPush.from_webdump(tmp_path, 6000)
    b = U.PageRankPush(rp, col, 6000)
    try:
        a.run(4, U.DEVICE_CPU)
        b.run(4, U.DEVICE_CPU)
        ya, yb = a.ranks(), b.ranks()
    finally:
        a.close()
        b.close()
    assert _rel_l1(ya, O.pagerank(rp, col, 4, 6000)) <= 1e-5
    assert _rel_l1(ya, yb) <= 1e-6
