"""Writes the golden vectors under tests/golden/ from the reference's own serial functions.

The reference ships no fixtures (its tests are runtime differential checks with random
time-seeded inputs), so these vectors are produced here: small seeded inputs are fed to
oracle/_ref/ref_runner, which is built by oracle/Makefile from the reference's unmodified
sources (serialImageFilter, serialPageRank, int serialMatrixMultiply, serialMatrixTranspose,
transposeMatrixBlock). Run in a container that
has the reference sources:  make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle_lib as O  # noqa: E402


def csr_from_edges(n, edges):
    rp = np.zeros(n + 1, dtype=np.uint32)
    for s, _ in edges:
        rp[s + 1] += 1
    rp = np.cumsum(rp).astype(np.uint32)
    col = np.zeros(len(edges), dtype=np.uint32)
    pos = rp[:-1].astype(np.int64).copy()
    for s, t in edges:
        col[pos[s]] = t
        pos[s] += 1
    return rp, col


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/ref_runner missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20240607)
    # imageFiltering: serialImageFilter (char arithmetic), inverted input rows
    for name, (H, W, r) in {"if_a": (23, 37, 1), "if_b": (40, 33, 2), "if_c": (6, 5, 3), "if_d": (1, 9, 2)}.items():
        img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
        out = O.ref_imagefilter(img, w, r)
        np.savez_compressed(HERE / f"{name}.npz", image=img, weights=w, radius=np.int32(r), out=out)
    # pageRank: serialPageRank (2 iterations) over padded 1-based records
    for name, (n, deg_max, p_dangle) in {"pr_a": (64, 6, 0.1), "pr_b": (400, 20, 0.05)}.items():
        edges = []
        for s in range(n):
            if rng.random() < p_dangle:
                continue
            k = int(rng.integers(1, deg_max + 1))
            ts = rng.choice(n, size=min(k, n - 1), replace=False)
            edges += [(s, int(t)) for t in ts if t != s]
        rp, col = csr_from_edges(n, edges)
        max_out = int(np.diff(rp.astype(np.int64)).max())
        y = O.ref_pagerank_2iter(rp, col, max_out)
        np.savez_compressed(HERE / f"{name}.npz", row_ptr=rp, col=col, ranks=y)
    # matrixMultiply (int): serialMatrixMultiply
    for name, n in {"mm_a": 16, "mm_b": 33}.items():
        a = rng.integers(-16, 17, (n, n)).astype(np.int32)
        b = rng.integers(-16, 17, (n, n)).astype(np.int32)
        c = O.ref_matmul_int(a, b)
        np.savez_compressed(HERE / f"{name}.npz", a=a, b=b, c=c)
    # matrixTranspose: serialMatrixTranspose (both forms) and transposeMatrixBlock (own seed, so the
    # vectors above stay as they were)
    rt = np.random.default_rng(20250114)
    for name, (rows, cols, inplace) in {"mt_a": (6, 10, False), "mt_b": (12, 4, True), "mt_c": (1, 7, False),
                                         "mt_d": (16, 16, True)}.items():
        m = rt.integers(-(1 << 31), 1 << 31, (rows, cols), dtype=np.int64).astype(np.int32)
        np.savez_compressed(HERE / f"{name}.npz", a=m, inplace=np.int32(inplace), out=O.ref_transpose(m, inplace))
    blk = rt.integers(-1000, 1000, (9, 9)).astype(np.int32)
    np.savez_compressed(HERE / "mt_block.npz", a=blk, out_copy=O.ref_transpose_block(blk, False),
                        out_inplace=O.ref_transpose_block(blk, True))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
