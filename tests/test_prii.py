"""pageRankInvertedIndex on the CPU: the oracle pinned by independent formulations, and the app's
CPU path (invert task with data redistribution, then the pageRank tasks) against it.

Pins:
  - or_prii_invert (invertIndex_cpu + redistribution, reference pageRankInvertedIndex.cpp:301-399)
    against a numpy restatement by sorting: every link of source subtask s gets the key
    (group, s, -enumeration index); a stable sort by that key is the redistributed index;
  - invariants: the index holds exactly the dump's (src, dest) pairs with the source's out-degree,
    and group g holds exactly the links into pages [g P, (g + 1) P);
  - or_prii_pagerank against the FP64 matrix form y = 0.85 M x (M[d][s] = links s->d / outdeg s)
    within 1e-6 relative L1, and against closed forms (a directed cycle: 0.85 then 0.7225; all
    pages linking to page 0: page 0 gets 0.85 n, the rest 0);
  - or_prii_serial (serialPageRank, 208-247) within 1e-6 relative L1 of the subtask order.
The app's CPU path must equal the oracle bit for bit: the index (integer data) and the ranks
(same float operations in the same order).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O


def _graph(pages, seed, mean_out=6.0, max_out=20, dangling=0.1):
    import unicorn_amd as U
    return U.webgen(pages, seed, mean_out, max_out, dangling)


def _sorted_index(records, P):
    pages, rec = records.shape
    keys, rows = [], []
    for p in range(pages):
        c = int(records[p, 0])
        s = p // P
        for k in range(c):
            t = int(records[p, 1 + k])
            keys.append(((t - 1) // P, s, -(len(rows))))
            rows.append((p, t - 1, c))
    order = sorted(range(len(rows)), key=lambda i: keys[i])
    return np.array([rows[i] for i in order], np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("pages,P,seed", [(100, 30, 1), (1000, 250, 2), (2000, 2000, 3), (1537, 100, 4)])
def test_oracle_invert_matches_sort_form(pages, P, seed):
    row_ptr, col = _graph(pages, seed)
    rec = O.prii_records(row_ptr, col, 20)
    e, counts = O.prii_invert(rec, P)
    np.testing.assert_array_equal(e, _sorted_index(rec, P))
    # invariants
    src = np.repeat(np.arange(pages), np.diff(row_ptr.astype(np.int64)))
    want = sorted(zip(src.tolist(), col.tolist()))
    assert sorted(zip(e[:, 0].tolist(), e[:, 1].tolist())) == want
    assert np.array_equal(e[:, 2], np.diff(row_ptr.astype(np.int64))[e[:, 0]])
    start = 0
    for g, c in enumerate(counts):
        d = e[start:start + c, 1]
        assert np.all((d >= g * P) & (d < (g + 1) * P))
        start += c
    assert start == e.shape[0]


def _fp64(row_ptr, col, iterations):
    pages = row_ptr.size - 1
    deg = np.diff(row_ptr.astype(np.int64))
    src = np.repeat(np.arange(pages), deg)
    x = np.ones(pages)
    for _ in range(iterations):
        y = np.zeros(pages)
        np.add.at(y, col, 0.85 * x[src] / deg[src])
        x = y
    return x


@pytest.mark.parametrize("pages,P", [(3000, 700), (5000, 5000)])
def test_oracle_pagerank_fp64_and_serial(pages, P):
    row_ptr, col = _graph(pages, 7)
    rec = O.prii_records(row_ptr, col, 20)
    e, counts = O.prii_invert(rec, P)
    for it in (1, 2, 3):
        got = O.prii_pagerank(e, counts, pages, P, it).astype(np.float64)
        ref = _fp64(row_ptr, col, it)
        assert np.abs(got - ref).sum() / np.abs(ref).sum() < 1e-6
        ser = O.prii_serial(rec, it).astype(np.float64)
        assert np.abs(ser - got).sum() / np.abs(got).sum() < 1e-6


def test_oracle_pagerank_closed_forms():
    n = 64
    row_ptr = np.arange(n + 1, dtype=np.uint32)
    cyc = ((np.arange(n) + 1) % n).astype(np.uint32)                # page p -> p + 1
    rec = O.prii_records(row_ptr, cyc, 4)
    e, counts = O.prii_invert(rec, 16)
    assert np.all(O.prii_pagerank(e, counts, n, 16, 1) == np.float32(0.85))
    assert np.allclose(O.prii_pagerank(e, counts, n, 16, 2), 0.85 * 0.85, rtol=1e-7)
    star = np.zeros(n, np.uint32)                                    # every page -> page 0
    e, counts = O.prii_invert(O.prii_records(row_ptr, star, 4), 16)
    r = O.prii_pagerank(e, counts, n, 16, 1)
    assert abs(r[0] - 0.85 * n) < 1e-4 and np.all(r[1:] == 0)


def test_planted_order_error_is_caught():
    """Filling the groups upwards instead of downwards is a plausible slip: the sort form sees it."""
    row_ptr, col = _graph(300, 9)
    rec = O.prii_records(row_ptr, col, 20)
    e, counts = O.prii_invert(rec, 100)
    wrong, start = e.copy(), 0
    # reverse each (group, source subtask) run
    for g, c in enumerate(counts):
        blk = wrong[start:start + c]
        s = blk[:, 0] // 100
        for v in np.unique(s):
            idx = np.nonzero(s == v)[0]
            blk[idx] = blk[idx[::-1]]
        start += c
    assert not np.array_equal(wrong, _sorted_index(rec, 100))


def _dump(tmp_path, pages, seed, per_file, max_out=20, **kw):
    import unicorn_amd as U
    row_ptr, col = _graph(pages, seed, max_out=max_out, **kw)
    d = tmp_path / f"dump_{pages}_{seed}"
    U.webdump_write(str(d), row_ptr, col, max_out, per_file)
    return d, row_ptr, col, O.prii_records(row_ptr, col, max_out)


def _run(U, d, P, device, iterations=2):
    a = U.PageRankInvertedIndex(str(d), P)
    try:
        a.run(device, iterations)
        e, counts = a.index()
        return a, e, counts, a.ranks()
    finally:
        a.close()


@pytest.mark.parametrize("pages,per_file,P", [(2000, 100, 500), (3001, 250, 750), (1000, 1000, 0), (4096, 64, 100)])
def test_app_cpu_matches_oracle(runtime, tmp_path, pages, per_file, P):
    d, row_ptr, col, rec = _dump(tmp_path, pages, 11, per_file)
    a, e, counts, ranks = _run(runtime, d, P, runtime.DEVICE_CPU)
    Peff = a.pages_per_subtask
    eo, co = O.prii_invert(rec, Peff)
    np.testing.assert_array_equal(counts, co)
    np.testing.assert_array_equal(e, eo)
    np.testing.assert_array_equal(ranks, O.prii_pagerank(eo, co, pages, Peff, 2))


def test_app_pages_per_subtask_rounding(runtime, tmp_path):
    """DoInit's rounding: whole files per subtask, at least one file."""
    import unicorn_amd as U
    d, *_ = _dump(tmp_path, 1000, 3, 300)
    for want, got in ((0, 9999900), (700, 600), (200, 300), (900, 900)):
        a = U.PageRankInvertedIndex(str(d), want)
        assert a.pages_per_subtask == got
        a.close()


def test_app_group_without_inlinks(runtime, tmp_path):
    """A subtask whose pages receive no links still runs (zero inlinks, zero ranks)."""
    import unicorn_amd as U
    n, P = 400, 100
    row_ptr = np.arange(n + 1, dtype=np.uint32)
    col = (np.arange(n) % 100).astype(np.uint32)                 # every link lands in pages 0..99
    d = tmp_path / "skew"
    U.webdump_write(str(d), row_ptr, col, 4, 100)
    a, e, counts, ranks = _run(runtime, d, P, runtime.DEVICE_CPU, 1)
    assert list(counts) == [n, 0, 0, 0]
    assert np.all(ranks[100:] == 0) and np.allclose(ranks[:100], 0.85 * 4)


def test_app_rejects_malformed_links(runtime, tmp_path):
    import unicorn_amd as U
    n = 200
    row_ptr = np.arange(n + 1, dtype=np.uint32)
    col = np.zeros(n, np.uint32)
    d = tmp_path / "bad"
    U.webdump_write(str(d), row_ptr, col, 2, 100)
    f = d / "web" / "page_1"
    raw = np.fromfile(f, np.uint32)
    raw[1] = n + 5                                                # a target beyond the last page
    raw.tofile(f)
    a = U.PageRankInvertedIndex(str(d), 100)
    try:
        with pytest.raises(RuntimeError):
            a.run(U.DEVICE_CPU)
    finally:
        a.close()
