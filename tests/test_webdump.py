"""The reference's on-disk web dump (pageRank.cpp:53-197, scripts/webGenerator.pl:121-274).

A dump is a metadata file of three uint32 (pages, max out-links, pages per file) and files
web/page_<n> of fixed records: out-link count, that many 1-based target ids, zero padding to the
maximum. The hand-made dump below is written byte by byte from that description (independently
of the library's writer), with a ragged last file, padding, a duplicate link and a self link;
the reader must return exactly its out-link CSR. The writer must round-trip a generated graph,
and pageRank over a loaded dump must match the oracle on the same CSR.
"""
import struct

import numpy as np
import pytest

from oracle import oracle_lib as O


@pytest.fixture(scope="module")
def U():
    import unicorn_amd as U
    return U


def _write_raw(d, pages, max_out, per_file, out_links):
    """out_links[p] = 1-based targets of page p + 1."""
    (d / "web").mkdir(parents=True, exist_ok=True)
    (d / "web_dump_metadata").write_bytes(struct.pack("<3I", pages, max_out, per_file))
    for first in range(0, pages, per_file):
        recs = b""
        for p in range(first, min(pages, first + per_file)):
            links = out_links[p]
            recs += struct.pack(f"<{1 + max_out}I", len(links), *links, *([0] * (max_out - len(links))))
        (d / "web" / f"page_{first + 1}").write_bytes(recs)


def test_hand_made_dump(U, tmp_path):
    out_links = [[2, 3, 2], [], [3, 5], [1], [4, 1, 2]]   # page 1 links to 2 twice; page 3 to itself
    _write_raw(tmp_path, 5, 3, 2, out_links)
    assert U.webdump_meta(tmp_path) == (5, 3, 2)
    rp, col = U.webdump_read(tmp_path)
    np.testing.assert_array_equal(rp, [0, 3, 3, 5, 6, 9])
    np.testing.assert_array_equal(col, [1, 2, 1, 2, 4, 0, 3, 0, 1])


def test_round_trip(U, tmp_path):
    rp, col = U.webgen(20_000, 3)
    max_out = int(np.diff(rp.astype(np.int64)).max())
    U.webdump_write(tmp_path, rp, col, max_out, 3000)      # 7 files, the last one partial
    assert sorted(p.name for p in (tmp_path / "web").iterdir()) == sorted(
        f"page_{1 + 3000 * i}" for i in range(7))
    assert (tmp_path / "web" / "page_18001").stat().st_size == 2000 * (max_out + 1) * 4
    rp2, col2 = U.webdump_read(tmp_path)
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(col2, col)


@pytest.mark.parametrize("case", ["target_too_big", "count_too_big", "missing_file", "short_file"])
def test_invalid_dumps(U, tmp_path, case):
    out_links = [[2], [1, 3], [1]]
    if case == "target_too_big":
        out_links[1] = [1, 4]
    _write_raw(tmp_path, 3, 2, 2, out_links)
    if case == "count_too_big":
        b = bytearray((tmp_path / "web" / "page_1").read_bytes())
        b[0:4] = struct.pack("<I", 3)
        (tmp_path / "web" / "page_1").write_bytes(bytes(b))
    if case == "missing_file":
        (tmp_path / "web" / "page_3").unlink()
    if case == "short_file":
        (tmp_path / "web" / "page_3").write_bytes(b"\x01\x00")
    with pytest.raises(RuntimeError, match="web dump"):
        U.webdump_read(tmp_path)


def test_pagerank_from_dump_on_cpu(runtime, tmp_path):
    U = runtime
    rp, col = U.webgen(30_000, 8)
    U.webdump_write(tmp_path, rp, col, 100, 4096)
    pr = U.PageRank.from_webdump(tmp_path, 7_000)
    try:
        pr.run(4, U.DEVICE_CPU)
        y = pr.ranks()
    finally:
        pr.close()
    want = O.pagerank(rp, col, 4, 7_000)
    assert float(np.abs(y.astype(np.float64) - want).sum() / np.abs(want).sum()) <= 1e-5
