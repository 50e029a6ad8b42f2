"""The pageRank graph address space in its source-blocked layout (host code, no GPU).

The layout is a reordering of the in-link CSR, so it is pinned against the plain CSR built from
the same out-links: every row keeps exactly its in-links, each (row group, source block) segment
holds only sources of that block, bseg and bcnt agree with the segments, and the group
boundaries coincide with the CSR row pointers. A page with more than 255 in-links from one source
block cannot be held by the u8 counts and falls back to the plain CSR.
"""
import numpy as np
import pytest

GR, NB = 2048, 8                     # UH_PR_GROUP_ROWS, UH_PR_BLOCKS


@pytest.fixture(scope="module")
def U():
    import unicorn_amd as U
    return U


def _check_layout(U, rp_out, col_out):
    n = rp_out.size - 1
    rp, deg, icol, bseg, bcnt = U.pr_graph_layout(rp_out, col_out, blocked=True)
    crp, cdeg, ccol, cbseg, cbcnt = U.pr_graph_layout(rp_out, col_out, blocked=False)
    assert cbseg is None and cbcnt is None and bseg is not None
    np.testing.assert_array_equal(rp, crp)
    np.testing.assert_array_equal(deg, np.diff(rp_out.astype(np.int64)))
    np.testing.assert_array_equal(deg, cdeg)
    e = int(rp_out[-1])
    groups = (n + GR - 1) // GR
    bs = ((n + NB - 1) // NB + 31) // 32 * 32
    assert bseg.size == groups * NB + 1 and bcnt.size == groups * NB * GR
    assert bseg[-1] == e and np.all(np.diff(bseg.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(bseg[0:groups * NB:NB], rp[0:n:GR][:groups])
    # plain CSR: sources ascending within each row, and they are the out-links reversed
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp_out.astype(np.int64)))
    order = np.lexsort((src, col_out[:e].astype(np.int64)))
    np.testing.assert_array_equal(ccol, src[order])
    counts = bcnt.reshape(groups, NB, GR).astype(np.int64)
    # per (row, block) counts from the CSR
    dst = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    want = np.zeros((groups * GR, NB), dtype=np.int64)
    np.add.at(want, (dst, ccol.astype(np.int64) // bs), 1)
    np.testing.assert_array_equal(counts.transpose(0, 2, 1).reshape(-1, NB), want)
    # each segment: its count total, its sources in block x, rows in order, sources ascending
    for g in range(groups):
        for x in range(NB):
            s0, s1 = int(bseg[g * NB + x]), int(bseg[g * NB + x + 1])
            c = counts[g, x]
            assert s1 - s0 == int(c.sum())
            seg = icol[s0:s1].astype(np.int64)
            assert np.all(seg // bs == x)
            rows = np.repeat(np.arange(GR), c)
            key = rows * (n + 1) + seg
            assert np.all(np.diff(key) > 0)
    # every row keeps exactly its in-links
    blk_rows = np.concatenate([np.repeat(g * GR + np.arange(GR), counts[g, x])
                               for g in range(groups) for x in range(NB)]) if e else np.zeros(0, np.int64)
    got = np.lexsort((icol.astype(np.int64), blk_rows))
    np.testing.assert_array_equal(icol[got], ccol)


def test_blocked_layout_webgen(U):
    rp, col = U.webgen(20_000, 7)
    _check_layout(U, rp, col)


def test_blocked_layout_ragged_and_tiny(U):
    rp, col = U.webgen(5_000, 3)           # last group partial, block size rounded to 32 pages
    _check_layout(U, rp, col)
    rp = np.array([0, 1, 2, 2], dtype=np.uint32)   # 3 pages, one dangling
    col = np.array([1, 0], dtype=np.uint32)
    _check_layout(U, rp, col)


def test_blocked_layout_overflow_falls_back(U):
    n = 6000                                # page 0 gets 5999 in-links, >255 from one block
    rp = np.arange(n + 1, dtype=np.uint32)
    rp[1:] -= 1
    rp[0] = 0
    col = np.zeros(n - 1, dtype=np.uint32)
    _, _, _, bseg, bcnt = U.pr_graph_layout(rp, col, blocked=True)
    assert bseg is None and bcnt is None


def test_blocked_layout_rejects_2_to_30_pages(U):
    """The blocked kernels address the rank vector with 32-bit byte offsets, so a graph of 2^30
    pages or more keeps the plain CSR (the builder answers before reading any array)."""
    L = U.apps_lib()
    dummy = np.zeros(4, dtype=np.uint32)
    bcnt = np.zeros(4, dtype=np.uint8)
    for pages in (1 << 30, (1 << 32) - 1):
        assert L.ua_pr_block_counts(pages, dummy.ctypes.data, dummy.ctypes.data, bcnt.ctypes.data) == 0
