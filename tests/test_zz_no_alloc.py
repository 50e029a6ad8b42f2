"""The allocation-free pageRank pull (uh_pagerank_pull_noalloc) against uh_pagerank_pull.

Both launch the same source-blocked kernels; the variant takes the partial sums from a static
device array instead of a buffer it allocates. Both add in a fixed order, so y and w_next must be
bitwise equal, including the rows outside the range (left untouched). Two noalloc calls run back
to back, the second over a smaller, unaligned row range, so the second reads partial sums over
the first's leftovers; the first covers the array's full capacity (512 row groups). Sizes past
the capacity and batch handles are refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GROUP_ROWS = 2048
MAX_GROUPS = 512
UH_ERR_ARG, UH_ERR_BATCH = 2, 5
SENTINEL = 0x7FC00001            # a quiet NaN: rows outside the range keep it


class _Graph:
    def __init__(self, U, pages, seed):
        from unicorn_amd import hip as H
        self.H = H
        rp, col = U.webgen(pages, seed)
        in_rp, deg, icol, bseg, bcnt = U.pr_graph_layout(rp, col, blocked=True)
        assert bcnt is not None, "graph did not take the source-blocked layout"
        self.pages = pages
        rng = np.random.default_rng(seed)
        self.w = rng.random(pages, dtype=np.float32)
        # y of every page in float64, from the out-link CSR (independent of the layout)
        self.ref = np.bincount(col.astype(np.int64), weights=np.repeat(self.w.astype(np.float64), np.diff(rp)),
                               minlength=pages)
        self.buf = {k: H.DeviceBuffer.from_array(np.ascontiguousarray(v)) for k, v in
                    (("rp", in_rp), ("col", icol), ("deg", deg), ("bseg", bseg), ("bcnt", bcnt), ("w", self.w))}
        for k in ("yA", "wA", "yB", "wB"):
            self.buf[k] = H.DeviceBuffer(4 * pages)

    def args(self, y, wn, row0, rows):
        b = self.buf
        return self.H.uh_pagerank_args(b["rp"].ptr, b["col"].ptr, b["deg"].ptr, b["w"].ptr, b[y].ptr, b[wn].ptr,
                                       row0, rows, 0.85, 0, b["bcnt"].ptr, b["bseg"].ptr)

    def reset(self):
        L = self.H.lib()
        for k in ("yA", "wA", "yB", "wB"):
            self.H.check(L.uh_fill_u32(self.buf[k].ptr, SENTINEL, self.pages, None), "fill")
        self.H.check(L.uh_device_sync(), "sync")

    def get(self, k):
        return self.buf[k].to_array(self.pages, np.uint32)

    def free(self):
        for v in self.buf.values():
            v.free()


def _run_pair(G, row0, rows, stream):
    H = G.H
    L = H.lib()
    G.reset()
    a = G.args("yA", "wA", row0, rows)
    H.check(L.uh_pagerank_pull(C.byref(a), stream), "uh_pagerank_pull")
    b = G.args("yB", "wB", row0, rows)
    H.check(L.uh_pagerank_pull_noalloc(C.byref(b), stream), "uh_pagerank_pull_noalloc")
    H.check(L.uh_stream_sync(stream), "sync")
    yA, yB, wA, wB = G.get("yA"), G.get("yB"), G.get("wA"), G.get("wB")
    assert np.array_equal(yA, yB) and np.array_equal(wA, wB)
    inside = np.zeros(G.pages, dtype=bool)
    inside[row0:row0 + rows] = True
    assert np.all(yB[~inside] == SENTINEL) and np.all(wB[~inside] == SENTINEL)
    # the sums themselves, against float64
    y = yB.view(np.float32)[inside].astype(np.float64)
    ref = G.ref[inside]
    assert np.abs(y - ref).sum() <= 1e-6 * max(np.abs(ref).sum(), 1e-30)
    return yB


def test_noalloc_matches_allocating_pull(gpu_runtime):
    from unicorn_amd import hip as H
    L = H.lib()
    pages = MAX_GROUPS * GROUP_ROWS + 3000
    G = _Graph(gpu_runtime, pages, 17)
    s = H.Stream()
    try:
        print(f"noalloc: {pages} pages, rows [0, {MAX_GROUPS * GROUP_ROWS})", flush=True)
        _run_pair(G, 0, MAX_GROUPS * GROUP_ROWS, s.h)           # the array's full capacity
        print("noalloc: rows [5000, 75000)", flush=True)
        _run_pair(G, 5000, 70_000, s.h)                          # smaller and unaligned, right after
        # sizes beyond the array and batch handles are refused before any launch
        big = G.args("yB", "wB", 1, MAX_GROUPS * GROUP_ROWS)     # touches 513 row groups
        assert L.uh_pagerank_pull_noalloc(C.byref(big), s.h) == UH_ERR_ARG
        assert b"row groups" in L.uh_last_error()
        bh = H.vp()
        H.check(L.uh_batch_create(C.byref(bh), s.h), "uh_batch_create")
        try:
            small = G.args("yB", "wB", 0, 100)
            assert L.uh_pagerank_pull_noalloc(C.byref(small), bh.value) == UH_ERR_BATCH
        finally:
            H.check(L.uh_batch_destroy(bh.value), "uh_batch_destroy")
    finally:
        s.sync()
        G.free()
