"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

1. Golden vectors written by the reference's own serial functions (tests/golden/*.npz, see
   make_golden.py): serialImageFilter, serialPageRank, int serialMatrixMultiply.
2. The reference's functions themselves (oracle/_ref/ref_runner) on fresh inputs, including its
   tile callback imageFilter_cpu, when the runner was built in this container.
3. Closed forms and invariants chosen so that plausible mistakes (flipped rows, swapped
   channels, transposed weights, dropped damping, wrong LU convention) fail one of them; a few
   tests plant such a mistake in the oracle's inputs and require the check to catch it.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle_lib as O

GOLD = Path(__file__).resolve().parent / "golden"


def _load(name):
    return np.load(GOLD / f"{name}.npz", allow_pickle=False)


# ---------------------------------------------------------------- image filter

@pytest.mark.parametrize("name", ["if_a", "if_b", "if_c", "if_d"])
def test_imagefilter_golden(name):
    g = _load(name)
    out = O.imagefilter(g["image"], g["weights"], int(g["radius"]), mode=0)
    assert np.array_equal(out, g["out"])


def test_imagefilter_golden_catches_planted_mistakes():
    g = _load("if_b")
    img, w, r = g["image"], g["weights"], int(g["radius"])
    D = 2 * r + 1
    wt = w.reshape(D, D).T.reshape(-1)
    assert not np.array_equal(O.imagefilter(img, wt, r, 0), g["out"])            # transposed weights
    assert not np.array_equal(O.imagefilter(img[::-1].copy(), w, r, 0), g["out"])  # rows not inverted
    assert not np.array_equal(O.imagefilter(img[:, :, ::-1].copy(), w, r, 0), g["out"])  # no channel swap


def test_imagefilter_identity_is_flip_and_channel_swap():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (11, 13, 3), dtype=np.uint8)
    for r in (1, 2):
        w = np.zeros((2 * r + 1) ** 2, dtype=np.int8)
        w[((2 * r + 1) ** 2) // 2] = 1
        want = img[::-1, :, ::-1]
        assert np.array_equal(O.imagefilter(img, w, r, 0), want)
        assert np.array_equal(O.imagefilter(img, w, r, 1, 0), want)


def test_imagefilter_constant_image():
    rng = np.random.default_rng(2)
    w = rng.integers(-128, 128, 25).astype(np.int8)
    for c in (0, 1, 77, 255):
        img = np.full((9, 7, 3), c, dtype=np.uint8)
        ref = (c * int(w.astype(np.int64).sum())) % 256      # (c * sum(w)) mod 256
        assert np.all(O.imagefilter(img, w, 2, 0) == ref)
    w4 = rng.integers(-4, 5, 25).astype(np.int8)
    img = np.full((9, 7, 3), 200, dtype=np.uint8)
    v = (200 * int(w4.astype(np.int64).sum())) >> 4
    assert np.all(O.imagefilter(img, w4, 2, 1, 4) == min(255, max(0, v)))


def test_imagefilter_delta_reproduces_weights():
    """A single bright pixel away from the borders spreads the (flipped) weights."""
    r, H, W = 2, 15, 15
    w = (np.arange(25, dtype=np.int64) - 12).astype(np.int8)
    img = np.zeros((H, W, 3), dtype=np.uint8)
    img[H - 1 - 7, 7, 0] = 1          # image row 7 (stored inverted), channel 0
    out = O.imagefilter(img, w, r, 0)
    patch = out[5:10, 5:10, 2].astype(np.int8)   # channel 0 lands in output channel 2
    assert np.array_equal(patch, w.reshape(5, 5)[::-1, ::-1])


def test_imagefilter_brute_force_small():
    rng = np.random.default_rng(3)
    for (H, W, r) in [(5, 4, 1), (3, 7, 2), (6, 6, 3)]:
        img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
        D = 2 * r + 1
        want = np.zeros_like(img)
        for i in range(H):
            for j in range(W):
                acc = np.zeros(3, dtype=np.int64)
                for dy in range(-r, r + 1):
                    for dx in range(-r, r + 1):
                        m = min(max(i + dy, 0), H - 1)
                        n = min(max(j + dx, 0), W - 1)
                        acc += img[H - 1 - m, n].astype(np.int64) * int(w[(dy + r) * D + dx + r])
                want[i, j] = (acc[::-1] % 256).astype(np.uint8)
        assert np.array_equal(O.imagefilter(img, w, r, 0), want)


def _raw_sums_by_padding(img_inv, w, r):
    """Per-pixel int64 sums of the filter, computed by edge-padding the un-inverted image and
    adding shifted slices (a second, vectorised formulation: no per-pixel clamp loop)."""
    H, W, _ = img_inv.shape
    D = 2 * r + 1
    up = img_inv[::-1].astype(np.int64)                     # image rows in top-down order
    pad = np.pad(up, ((r, r), (r, r), (0, 0)), mode="edge")
    acc = np.zeros((H, W, 3), dtype=np.int64)
    for dy in range(D):
        for dx in range(D):
            acc += int(w[dy * D + dx]) * pad[dy:dy + H, dx:dx + W]
    return acc[:, :, ::-1]                                  # output channel c sums input 2 - c


def test_imagefilter_mode1_brute_force_random():
    """Shift + clamp mode on random images against the padded-slice formulation: floor shift of
    negative sums (Python's >> floors), both clamp ends, radii 1..3, several shifts."""
    rng = np.random.default_rng(31)
    below = above = 0
    for (H, W, r, lo, hi, shift) in [(9, 11, 1, -4, 5, 4), (17, 13, 2, -4, 5, 4), (12, 20, 2, -4, 5, 0),
                                      (8, 8, 3, -128, 128, 7), (31, 5, 2, -16, 17, 2), (5, 40, 1, -4, 5, 1)]:
        img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
        w = rng.integers(lo, hi, (2 * r + 1) ** 2).astype(np.int8)
        acc = _raw_sums_by_padding(img, w, r)
        want = np.clip(acc >> shift, 0, 255).astype(np.uint8)
        got = O.imagefilter(img, w, r, 1, shift)
        assert np.array_equal(got, want), (H, W, r, shift)
        below += int((acc < 0).sum())
        above += int((acc >> shift > 255).sum())
        if shift:
            trunc = np.clip(np.sign(acc) * (np.abs(acc) >> shift), 0, 255)
            assert np.array_equal(trunc, want)  # truncation differs only below zero, clamped away
            wrapped = ((acc >> shift) % 256).astype(np.uint8)
            assert not np.array_equal(wrapped, want)                 # no clamp: caught
    assert below > 0 and above > 0                              # both clamp ends exercised
    # mode 0 against the same formulation (mod 256)
    img = rng.integers(0, 256, (23, 19, 3), dtype=np.uint8)
    w = rng.integers(-128, 128, 25).astype(np.int8)
    assert np.array_equal(O.imagefilter(img, w, 2, 0), (_raw_sums_by_padding(img, w, 2) % 256).astype(np.uint8))


def test_imagefilter_mode1_brute_force_catches_planted_mistakes():
    rng = np.random.default_rng(32)
    img = rng.integers(0, 256, (16, 16, 3), dtype=np.uint8)
    w = rng.integers(-4, 5, 25).astype(np.int8)
    want = np.clip(_raw_sums_by_padding(img, w, 2) >> 4, 0, 255).astype(np.uint8)
    wt = w.reshape(5, 5).T.reshape(-1)
    assert not np.array_equal(O.imagefilter(img, wt, 2, 1, 4), want)                  # transposed weights
    assert not np.array_equal(O.imagefilter(img, w, 2, 1, 3), want)                   # wrong shift
    assert not np.array_equal(O.imagefilter(img[::-1].copy(), w, 2, 1, 4), want)      # rows not inverted


def test_imagefilter_tiles_equal_whole_image():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (70, 90, 3), dtype=np.uint8)
    w = rng.integers(-4, 5, 25).astype(np.int8)
    for tile in (16, 32, 64):
        ntiles = ((90 + tile - 1) // tile) * ((70 + tile - 1) // tile)
        assert np.array_equal(O.imagefilter_tiled(img, w, 2, 1, 4, tile, 0, ntiles, 3), O.imagefilter(img, w, 2, 1, 4))


@pytest.mark.skipif(not O.ref_available(), reason="reference runner not built in this environment")
def test_imagefilter_reference_tile_callback():
    """imageFilter_cpu of the reference over its 2048^2 tiles equals the oracle's tiling."""
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (2100, 4200, 3), dtype=np.uint8)
    w = rng.integers(-128, 128, 25).astype(np.int8)
    ref = O.ref_imagefilter_tiles(img, w, 2)
    assert np.array_equal(O.imagefilter_tiled(img, w, 2, 0, 4, 2048, 0, 6, 8), ref)


# ---------------------------------------------------------------- pageRank

@pytest.mark.parametrize("name", ["pr_a", "pr_b"])
def test_pagerank_golden(name):
    g = _load(name)
    y = O.pagerank(g["row_ptr"], g["col"], 2, 0)
    assert np.array_equal(y, g["ranks"])


def test_pagerank_golden_catches_planted_mistake():
    g = _load("pr_b")
    assert not np.array_equal(O.pagerank(g["row_ptr"], g["col"], 2, 0, damping=1.0), g["ranks"])
    assert not np.array_equal(O.pagerank(g["row_ptr"], g["col"], 1, 0), g["ranks"])


@pytest.mark.skipif(not O.ref_available(), reason="reference runner not built in this environment")
def test_pagerank_reference_fresh_graph():
    import unicorn_amd as U
    rp, col = U.webgen(3000, 19, 10.0, 100, 0.05)
    max_out = int(np.diff(rp.astype(np.int64)).max())
    assert np.array_equal(O.pagerank(rp, col, 2, 0), O.ref_pagerank_2iter(rp, col, max_out))


def test_pagerank_closed_forms():
    n = 50
    rp = np.arange(n + 1, dtype=np.uint32)
    col = ((np.arange(n) + 1) % n).astype(np.uint32)          # ring
    for it in (1, 2, 5):
        np.testing.assert_allclose(O.pagerank(rp, col, it, 7), np.float32(0.85) ** it, rtol=1e-6)
    # star into page 0 from every other page
    rp = np.concatenate([[0, 0], np.arange(1, n, dtype=np.int64)]).astype(np.uint32)
    col = np.zeros(n - 1, dtype=np.uint32)
    y = O.pagerank(rp, col, 1, 16)
    assert abs(float(y[0]) - 0.85 * (n - 1)) <= 1e-6 * 0.85 * (n - 1) and np.all(y[1:] == 0)
    # all dangling
    assert np.all(O.pagerank(np.zeros(n + 1, dtype=np.uint32), np.zeros(0, dtype=np.uint32), 3, 9) == 0)


def test_pagerank_mass_invariant():
    """sum(y) = 0.85 * sum over non-dangling sources of x (up to rounding)."""
    import unicorn_amd as U
    rp, col = U.webgen(5000, 23)
    deg = np.diff(rp.astype(np.int64))
    x1 = O.pagerank_f64(rp, col, 1)
    x2 = O.pagerank_f64(rp, col, 2)
    assert abs(x2.sum() - 0.85 * x1[deg > 0].sum()) <= 1e-9 * x2.sum()
    y = O.pagerank(rp, col, 2, 1000)
    assert np.abs(y - x2).sum() / x2.sum() <= 1e-6


# ---------------------------------------------------------------- matrix multiply

@pytest.mark.parametrize("name", ["mm_a", "mm_b"])
def test_matmul_golden_int(name):
    g = _load(name)
    n = g["a"].shape[0]
    got = O.matmul_blocks(g["a"].astype(np.float32), g["b"].astype(np.float32), n)   # exact for |v| <= 16
    assert np.array_equal(got.astype(np.int64), g["c"].astype(np.int64))
    got64 = O.matmul_f64(g["a"].astype(np.float32), g["b"].astype(np.float32))
    assert np.array_equal(got64.astype(np.int64), g["c"].astype(np.int64))


def test_matmul_block_decomposition():
    rng = np.random.default_rng(6)
    A = rng.integers(-8, 9, (64, 64)).astype(np.float32)
    B = rng.integers(-8, 9, (64, 64)).astype(np.float32)
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    for bs in (8, 16, 32, 64):
        assert np.array_equal(O.matmul_blocks(A, B, bs, threads=3), want)
    # a planted transpose is caught
    assert not np.array_equal(O.matmul_blocks(A, B.T.copy(), 16), want)


# ---------------------------------------------------------------- LU

def test_lu_reconstructs_constructed_factors():
    """A = L U with L lower (pivots on the diagonal) and U unit upper: factoring A returns them."""
    rng = np.random.default_rng(8)
    n = 24
    L = np.tril(rng.integers(-3, 4, (n, n))).astype(np.float64)
    np.fill_diagonal(L, rng.integers(1, 4, n) * rng.choice([-1, 1], n))
    Uu = np.triu(rng.integers(-3, 4, (n, n)), 1).astype(np.float64) + np.eye(n)
    A = L @ Uu
    F = O.lu_f64(np.asfortranarray(A))
    np.testing.assert_allclose(np.tril(F), L, atol=1e-9)
    np.testing.assert_allclose(np.triu(F, 1) + np.eye(n), Uu, atol=1e-9)
    F32 = O.lu(np.asfortranarray(A.astype(np.float32)))
    np.testing.assert_allclose(np.tril(F32), L, rtol=1e-4, atol=1e-3)


def test_lu_diagonal_matrix():
    d = np.arange(1, 9, dtype=np.float64)
    F = O.lu_f64(np.asfortranarray(np.diag(d)))
    assert np.array_equal(F, np.diag(d))


def test_lu_solve_and_residual():
    rng = np.random.default_rng(9)
    n = 40
    A = (rng.random((n, n)) + n * np.eye(n)).astype(np.float32)
    x = rng.uniform(-1, 1, (n, 3))
    b = np.asfortranarray(A.astype(np.float64) @ x)
    F = O.lu_f64(np.asfortranarray(A.astype(np.float64)))
    xs = O.lu_solve_f64(F, b)
    assert np.abs(xs - x).max() <= 1e-12
    r = O.matvec_residual_f64(np.asfortranarray(A), xs, b)
    assert np.abs(r).max() <= 1e-11


# ---------------------------------------------------------------- matrix transpose

@pytest.mark.parametrize("name", ["mt_a", "mt_b", "mt_c", "mt_d"])
def test_transpose_golden(name):
    g = _load(name)
    np.testing.assert_array_equal(O.transpose(g["a"]), g["out"])
    # the fixture is the closed form too: serialMatrixTranspose's cycles move every element
    np.testing.assert_array_equal(g["out"], g["a"].T)


def test_transpose_block_golden():
    g = _load("mt_block")
    np.testing.assert_array_equal(g["out_copy"], g["a"].T)
    np.testing.assert_array_equal(g["out_inplace"], g["a"].T)


def test_transpose_index_map_catches_planted_mistake():
    # the map (i * cols) mod (n - 1) with rows in place of cols is another permutation
    m = np.arange(6 * 10, dtype=np.int32).reshape(6, 10)
    n = m.size
    wrong = np.array([m.reshape(-1)[(i * 6) % (n - 1)] for i in range(n - 1)] + [m.reshape(-1)[-1]], dtype=np.int32)
    assert not np.array_equal(wrong.reshape(10, 6), O.transpose(m))


@pytest.mark.parametrize("rows,cols,bs", [(8, 4, 4), (12, 12, 4), (6, 18, 6), (5, 5, 5)])
def test_transpose_blocks_equal_whole_matrix(rows, cols, bs):
    m = np.random.default_rng(rows * cols).integers(-999, 999, (rows, cols)).astype(np.int32)
    np.testing.assert_array_equal(O.transpose_blocks(m, bs), O.transpose(m))


def test_transpose_reference_fresh_inputs():
    if not O.ref_available():
        pytest.skip("oracle/_ref/ref_runner is built only where the reference sources are")
    rng = np.random.default_rng(77)
    for rows, cols in [(3, 11), (32, 8), (9, 9)]:
        m = rng.integers(-(1 << 31), 1 << 31, (rows, cols), dtype=np.int64).astype(np.int32)
        for inplace in (False, True):
            np.testing.assert_array_equal(O.ref_transpose(m, inplace), O.transpose(m))


# ---------------------------------------------------------------- prefix sum

def test_prefix_sum_closed_form_and_chunks():
    # serialPrefixSum over int wraps modulo 2^32: the closed form is a 64-bit cumsum cast back
    rng = np.random.default_rng(31)
    x = rng.integers(-(1 << 31), 1 << 31, 10007, dtype=np.int64).astype(np.int32)
    want = np.cumsum(x.astype(np.int64)).astype(np.int32)
    np.testing.assert_array_equal(O.prefix_sum(x), want)
    for chunk in (1, 7, 1000, 10007, 20000):
        np.testing.assert_array_equal(O.prefix_sum(x, chunk), want)


def test_prefix_sum_planted_mistakes():
    x = np.arange(1, 11, dtype=np.int32)
    assert O.prefix_sum(x)[-1] == 55
    exclusive = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32)
    assert not np.array_equal(O.prefix_sum(x), exclusive)
    # a chunked result without the carried totals differs after the first chunk
    no_carry = np.concatenate([np.cumsum(x[:4]), np.cumsum(x[4:8]), np.cumsum(x[8:])]).astype(np.int32)
    assert not np.array_equal(O.prefix_sum(x, 4), no_carry)


# ---------------------------------------------------------------- sparse x dense

def _coo(n, density, rng):
    k = max(1, int(n * n * density))
    idx = np.unique(rng.choice(n * n, size=k, replace=False))
    return (idx // n).astype(np.int32), (idx % n).astype(np.int32)


def test_spmm_coo_equals_dense_product():
    rng = np.random.default_rng(41)
    n = 64
    r, c = _coo(n, 0.05, rng)
    v = rng.integers(-9, 10, r.size).astype(np.float32)
    B = rng.integers(-9, 10, (n, n)).astype(np.float32)
    A = np.zeros((n, n), np.float64)
    A[r, c] = v
    want = A @ B.astype(np.float64)
    # small integers: float and double forms are both exact
    np.testing.assert_array_equal(O.spmm_coo(r, c, v, B, f64=True), want)
    np.testing.assert_array_equal(O.spmm_coo(r, c, v, B).astype(np.float64), want)


def test_spmm_coo_planted_mistake():
    rng = np.random.default_rng(42)
    n = 32
    r, c = _coo(n, 0.1, rng)
    v = rng.standard_normal(r.size).astype(np.float32)
    B = rng.standard_normal((n, n)).astype(np.float32)
    good = O.spmm_coo(r, c, v, B, f64=True)
    assert not np.allclose(O.spmm_coo(c, r, v, B, f64=True), good)          # A transposed
    assert np.allclose(O.spmm_coo(r, c, v, B), good, rtol=1e-5, atol=1e-5)



def test_spmm_float_order_brute_force():
    """The float form rounds each product, then each sum, entry by entry in COO order: a Python
    loop over np.float32 scalars gives the same bits; a transposed B is caught."""
    n = 12
    rng = np.random.default_rng(4)
    idx = np.unique(rng.choice(n * n, size=40, replace=False))
    r, c = (idx // n).astype(np.int32), (idx % n).astype(np.int32)
    v = rng.standard_normal(r.size).astype(np.float32)
    B = rng.standard_normal((n, n)).astype(np.float32)
    want = np.zeros((n, n), np.float32)
    for k in range(r.size):
        for j in range(n):
            want[r[k], j] = np.float32(want[r[k], j] + np.float32(v[k] * B[c[k], j]))
    np.testing.assert_array_equal(O.spmm_coo(r, c, v, B), want)
    assert not np.array_equal(O.spmm_coo(r, c, v, np.ascontiguousarray(B.T)), want)
