"""ORACLE — test infrastructure only (loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product).

ctypes bindings of oracle/liboracle.so (the C restatement in oracle.c) and helpers that drive
oracle/_ref/ref_runner (the reference's own serial functions, built in this container only).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_RUNNER = HERE / "_ref" / "ref_runner"

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle"], check=True)
    if Path("/root/reference").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle liboracle`")
        _lib = C.CDLL(str(LIB))
        vp, i32, u32, f32, f64 = C.c_void_p, C.c_int, C.c_uint32, C.c_float, C.c_double
        _lib.or_imagefilter.argtypes = [vp, i32, i32, vp, i32, i32, i32, vp]
        _lib.or_imagefilter_tiled.argtypes = [vp, i32, i32, vp, i32, i32, i32, i32, vp, i32, i32, i32]
        _lib.or_pagerank.argtypes = [u32, vp, vp, i32, u32, f32, vp]
        _lib.or_pagerank_f64.argtypes = [u32, vp, vp, i32, f64, vp]
        _lib.or_matmul_blocks.argtypes = [vp, vp, vp, i32, i32, i32, i32, i32]
        _lib.or_matmul_f64.argtypes = [vp, vp, vp, i32, i32, i32]
        _lib.or_lu.argtypes = [vp, i32, i32]
        _lib.or_lu_f64.argtypes = [vp, i32, i32]
        _lib.or_lu_solve_f64.argtypes = [vp, i32, i32, vp, i32]
        _lib.or_matvec_f64.argtypes = [vp, i32, i32, vp, vp, i32, vp]
        _lib.or_transpose_i32.argtypes = [vp, vp, C.c_size_t, C.c_size_t]
        _lib.or_transpose_blocks.argtypes = [vp, vp, C.c_size_t, C.c_size_t, C.c_size_t]
        _lib.or_prefix_sum_i32.argtypes = [vp, vp, C.c_size_t, C.c_size_t]
        _lib.or_spmm_coo.argtypes = [vp, vp, vp, C.c_size_t, vp, C.c_size_t, vp, vp, i32]
        _lib.or_fft2_f64.argtypes = [vp, vp, C.c_size_t, C.c_size_t]
        _lib.or_prii_invert.argtypes = [vp, u32, u32, u32, vp, vp]
        _lib.or_prii_invert.restype = C.c_uint64
        _lib.or_prii_pagerank.argtypes = [vp, vp, u32, u32, i32, vp]
        _lib.or_prii_serial.argtypes = [vp, u32, u32, i32, vp]
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data


def imagefilter(img_inv: np.ndarray, weights: np.ndarray, radius: int, mode: int = 0, shift: int = 4) -> np.ndarray:
    H, W, _ = img_inv.shape
    img_inv = np.ascontiguousarray(img_inv, dtype=np.uint8)
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.int8).reshape(-1))
    out = np.empty_like(img_inv)
    lib().or_imagefilter(_p(img_inv), W, H, _p(w), radius, mode, shift, _p(out))
    return out


def imagefilter_tiled(img_inv, weights, radius, mode, shift, tile, t0, t1, threads, out=None):
    H, W, _ = img_inv.shape
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.int8).reshape(-1))
    if out is None:
        out = np.zeros_like(img_inv)
    lib().or_imagefilter_tiled(_p(img_inv), W, H, _p(w), radius, mode, shift, tile, _p(out), t0, t1, threads)
    return out


def pagerank(row_ptr: np.ndarray, col: np.ndarray, iterations: int, subtask_pages: int = 0,
             damping: float = 0.85) -> np.ndarray:
    n = row_ptr.size - 1
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    out = np.empty(n, dtype=np.float32)
    lib().or_pagerank(n, _p(row_ptr), _p(col), iterations, subtask_pages, damping, _p(out))
    return out


def pagerank_f64(row_ptr, col, iterations, damping=0.85):
    n = row_ptr.size - 1
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    out = np.empty(n, dtype=np.float64)
    lib().or_pagerank_f64(n, _p(row_ptr), _p(col), iterations, damping, _p(out))
    return out


def matmul_blocks(A, B, bs, s0=None, s1=None, threads=1, C_out=None):
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    if C_out is None:
        C_out = np.zeros((n, n), dtype=np.float32)
    nb = (n // bs) ** 2
    lib().or_matmul_blocks(_p(A), _p(B), _p(C_out), n, bs, 0 if s0 is None else s0, nb if s1 is None else s1, threads)
    return C_out


def matmul_f64(A, B):
    M, K = A.shape
    N = B.shape[1]
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    Cm = np.empty((M, N), dtype=np.float64)
    lib().or_matmul_f64(_p(A), _p(B), _p(Cm), M, N, K)
    return Cm


def lu(a_colmajor: np.ndarray) -> np.ndarray:
    """a: n x n float32 in column-major order (a Fortran-ordered array)."""
    a = np.asfortranarray(a_colmajor, dtype=np.float32).copy(order="F")
    n = a.shape[0]
    lib().or_lu(_p(a), n, n)
    return a


def lu_f64(a_colmajor):
    a = np.asfortranarray(a_colmajor, dtype=np.float64).copy(order="F")
    n = a.shape[0]
    lib().or_lu_f64(_p(a), n, n)
    return a


def lu_solve_f64(lu_f, b):
    n = lu_f.shape[0]
    lu_f = np.asfortranarray(lu_f, dtype=np.float64)
    x = np.asfortranarray(b, dtype=np.float64).copy(order="F")
    nrhs = x.shape[1] if x.ndim == 2 else 1
    lib().or_lu_solve_f64(_p(lu_f), n, n, _p(x), nrhs)
    return x


def matvec_residual_f64(A_colmajor, x, b):
    n = A_colmajor.shape[0]
    A = np.asfortranarray(A_colmajor, dtype=np.float32)
    x = np.asfortranarray(x, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    nrhs = x.shape[1]
    r = np.empty((n, nrhs), dtype=np.float64, order="F")
    lib().or_matvec_f64(_p(A), n, n, _p(x), _p(b), nrhs, _p(r))
    return r


# ---------------------------------------------------------------- reference runner

def transpose(M: np.ndarray) -> np.ndarray:
    """serialMatrixTranspose's index map over a rows x cols int32 matrix."""
    M = np.ascontiguousarray(M, dtype=np.int32)
    rows, cols = M.shape
    out = np.empty((cols, rows), dtype=np.int32)
    lib().or_transpose_i32(_p(M), _p(out), rows, cols)
    return out


def transpose_blocks(M: np.ndarray, bs: int) -> np.ndarray:
    """The app's block decomposition (one bs x bs block per subtask)."""
    M = np.ascontiguousarray(M, dtype=np.int32)
    rows, cols = M.shape
    out = np.empty((cols, rows), dtype=np.int32)
    lib().or_transpose_blocks(_p(M), _p(out), rows, cols, bs)
    return out


def prefix_sum(x: np.ndarray, chunk: int = 0) -> np.ndarray:
    """serialPrefixSum (chunk 0), or the app's chunk scan + aux + add decomposition."""
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.empty_like(x)
    lib().or_prefix_sum_i32(_p(x), _p(out), x.size, chunk)
    return out


def spmm_coo(rows, cols, vals, B, f64=False):
    """C = A B for COO A (entries in order) and dense B: float with each product and sum rounded,
    or double."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    n = B.shape[0]
    Cf = np.zeros((n, n), np.float32)
    Cd = np.zeros((n, n), np.float64)
    lib().or_spmm_coo(_p(rows), _p(cols), _p(vals), rows.size, _p(B), n, _p(Cf), _p(Cd), int(f64))
    return Cd if f64 else Cf


def fft2(x: np.ndarray) -> np.ndarray:
    """fftSerial's forward 2-D DFT (rows, then columns) of a complex64 matrix, in double."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    out = np.empty(x.shape, np.complex128)
    lib().or_fft2_f64(_p(x), _p(out), x.shape[0], x.shape[1])
    return out


def prii_records(row_ptr: np.ndarray, col: np.ndarray, max_out: int) -> np.ndarray:
    """The web dump's padded records (pages x (max_out + 1) uint32) of an out-link CSR."""
    pages = row_ptr.size - 1
    rec = np.zeros((pages, max_out + 1), np.uint32)
    deg = np.diff(row_ptr.astype(np.int64))
    rec[:, 0] = deg
    src = np.repeat(np.arange(pages), deg)
    slot = np.arange(deg.sum()) - np.repeat(row_ptr[:-1].astype(np.int64), deg)
    rec[src, 1 + slot] = np.asarray(col[:deg.sum()], dtype=np.uint32) + 1
    return rec


def prii_invert(records: np.ndarray, P: int):
    """invertIndex_cpu per subtask + the redistribution: (entries [n x 3], counts per group)."""
    records = np.ascontiguousarray(records, dtype=np.uint32)
    pages, rec = records.shape
    links = int(records[:, 0].sum())
    entries = np.zeros((max(links, 1), 3), np.uint32)
    D = (pages + P - 1) // P
    counts = np.zeros(D, np.uint32)
    n = lib().or_prii_invert(_p(records), pages, rec - 1, P, _p(entries), _p(counts))
    return entries[:n], counts


def prii_pagerank(entries: np.ndarray, counts: np.ndarray, pages: int, P: int, iterations: int = 2) -> np.ndarray:
    """pageRank_cpu over the redistributed index, subtask by subtask, entries in order."""
    entries = np.ascontiguousarray(entries if entries.size else np.zeros((1, 3), np.uint32), dtype=np.uint32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    out = np.zeros(pages, np.float32)
    lib().or_prii_pagerank(_p(entries), _p(counts), pages, P, iterations, _p(out))
    return out


def prii_serial(records: np.ndarray, iterations: int = 2) -> np.ndarray:
    """serialPageRank of the inverted-index app over the padded records."""
    records = np.ascontiguousarray(records, dtype=np.uint32)
    out = np.zeros(records.shape[0], np.float32)
    lib().or_prii_serial(_p(records), records.shape[0], records.shape[1] - 1, iterations, _p(out))
    return out


def ref_available() -> bool:
    return REF_RUNNER.exists()


def _run_ref(args):
    subprocess.run([str(REF_RUNNER)] + [str(a) for a in args], check=True)


def ref_imagefilter(img_inv: np.ndarray, weights: np.ndarray, radius: int) -> np.ndarray:
    """serialImageFilter of the reference (char arithmetic)."""
    H, W, _ = img_inv.shape
    f = np.zeros((32, 32), dtype=np.int8)
    D = 2 * radius + 1
    f[:D, :D] = np.asarray(weights, dtype=np.int8).reshape(D, D)
    with tempfile.TemporaryDirectory() as d:
        pin, pf, po = Path(d, "in"), Path(d, "f"), Path(d, "o")
        np.ascontiguousarray(img_inv, dtype=np.uint8).tofile(pin)
        f.tofile(pf)
        _run_ref(["if", W, H, radius, pin, pf, po])
        return np.fromfile(po, dtype=np.uint8).reshape(H, W, 3)


def ref_imagefilter_tiles(img_inv, weights, radius):
    """imageFilter_cpu of the reference over its 2048x2048 tile subtasks (NATURAL views)."""
    H, W, _ = img_inv.shape
    f = np.zeros((32, 32), dtype=np.int8)
    D = 2 * radius + 1
    f[:D, :D] = np.asarray(weights, dtype=np.int8).reshape(D, D)
    with tempfile.TemporaryDirectory() as d:
        pin, pf, po = Path(d, "in"), Path(d, "f"), Path(d, "o")
        np.ascontiguousarray(img_inv, dtype=np.uint8).tofile(pin)
        f.tofile(pf)
        _run_ref(["iftile", W, H, radius, 2048, pin, pf, po])
        return np.fromfile(po, dtype=np.uint8).reshape(H, W, 3)


def ref_pagerank_2iter(row_ptr, col, max_out):
    """serialPageRank of the reference (2 iterations) over padded 1-based records."""
    n = row_ptr.size - 1
    rec = np.zeros((n, max_out + 1), dtype=np.uint32)
    deg = np.diff(row_ptr.astype(np.int64))
    rec[:, 0] = deg
    for p in range(n):
        rec[p, 1:1 + deg[p]] = col[row_ptr[p]:row_ptr[p + 1]] + 1
    with tempfile.TemporaryDirectory() as d:
        pr, po = Path(d, "r"), Path(d, "o")
        rec.tofile(pr)
        _run_ref(["pr", n, max_out, pr, po])
        return np.fromfile(po, dtype=np.float32)


def ref_matmul_int(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    n = A.shape[0]
    with tempfile.TemporaryDirectory() as d:
        pa, pb, pc = Path(d, "a"), Path(d, "b"), Path(d, "c")
        np.ascontiguousarray(A, dtype=np.int32).tofile(pa)
        np.ascontiguousarray(B, dtype=np.int32).tofile(pb)
        _run_ref(["mm", n, pa, pb, pc])
        return np.fromfile(pc, dtype=np.int32).reshape(n, n)


def ref_transpose(M: np.ndarray, inplace: bool) -> np.ndarray:
    """serialMatrixTranspose of the reference (out of place, or in place on a copy)."""
    rows, cols = M.shape
    with tempfile.TemporaryDirectory() as d:
        pa, po = Path(d, "a"), Path(d, "o")
        np.ascontiguousarray(M, dtype=np.int32).tofile(pa)
        _run_ref(["mt", int(inplace), rows, cols, pa, po])
        return np.fromfile(po, dtype=np.int32).reshape(cols, rows)


def ref_transpose_block(B: np.ndarray, inplace: bool) -> np.ndarray:
    """transposeMatrixBlock of the reference over one square block."""
    bs = B.shape[0]
    with tempfile.TemporaryDirectory() as d:
        pa, po = Path(d, "a"), Path(d, "o")
        np.ascontiguousarray(B, dtype=np.int32).tofile(pa)
        _run_ref(["mtblock", int(inplace), bs, pa, po])
        return np.fromfile(po, dtype=np.int32).reshape(bs, bs)
