"""Built-in subtask reductions and data redistribution on the CPU stubs (no GPU needed).

Reductions: every subtask of the probe task (tests/native/reduce_probe.hip) writes the whole
output space with small integers that all six element types hold exactly, and the task's
reduction is pmGetSubtaskReductionCallbackImpl(op, type) for each of the reference's ten
operations (pmReducer.cpp:593-703): ADD, MIN, MAX, PRODUCT, LOGICAL_AND / OR, LOGICAL_XOR
(a != b, as the reference writes it), BITWISE_AND / OR / XOR (floating values converted to
unsigned integers). The result must equal numpy's fold of the subtasks' values.

Redistribution: each subtask cuts its block into chunks of hashed lengths, passes the lengths to
its dataRedistribution callback in a SUBTASK_TO_POST_SUBTASK scratch buffer, and gives chunk j
the order (subtask + j) % orders. After the task the space must be the concatenation of all
chunks by (order, subtask), and pmGetRedistributionMetadata must list each order's bytes
(reference pmRedistributor.cpp, pmSubscriptionManager.cpp:861-877).
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

PROBE = Path(__file__).resolve().parent / "native" / "libreduce_probe.so"
OPS = ["ADD", "MIN", "MAX", "PRODUCT", "LOGICAL_AND", "BITWISE_AND", "LOGICAL_OR", "BITWISE_OR",
       "LOGICAL_XOR", "BITWISE_XOR"]
TYPES = [np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
M32 = 0xFFFFFFFF


def load(runtime):
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.rp_reduce.restype = C.c_int
    L.rp_reduce.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
    L.rp_reduce_ex.restype = C.c_int
    L.rp_reduce_ex.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_void_p]
    L.rp_redistribute.restype = C.c_int
    L.rp_redistribute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                  C.POINTER(C.c_uint32)]
    return L


@pytest.fixture(scope="module")
def lib(runtime):
    return load(runtime)


def _mix(a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    h = ((a * 0x9E3779B1) & M32) ^ ((((b + 0x7F4A7C15) & M32) * 0x85EBCA77) & M32)
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & M32
    h ^= h >> 12
    return h


def gen(op, t, sid, n, sparse=False):
    """Subtask sid's values (reduce_probe.hip gen); sparse keeps about one element in 16."""
    h = _mix(sid + 1, np.arange(n))
    uns = t in (np.uint32, np.uint64)
    name = OPS[op]
    if name == "PRODUCT":
        v = (h % 2 + 1).astype(np.int64)
    elif name in ("ADD", "MIN", "MAX"):
        v = (h % 7).astype(np.int64) - (0 if uns else 3)
    elif name == "LOGICAL_XOR":
        v = (h % 2).astype(np.int64)
    else:
        v = (h % 5).astype(np.int64)
    if sparse:
        v = np.where((_mix(sid + 101, np.arange(n)) & 15) != 0, 0, v)
    return v.astype(t)


def fold(op, t, vals):
    name = OPS[op]
    acc = vals[0].copy()
    for b in vals[1:]:
        if name == "ADD":
            acc = (acc + b).astype(t)
        elif name == "MIN":
            acc = np.minimum(acc, b)
        elif name == "MAX":
            acc = np.maximum(acc, b)
        elif name == "PRODUCT":
            acc = (acc * b).astype(t)
        elif name == "LOGICAL_AND":
            acc = ((acc != 0) & (b != 0)).astype(t)
        elif name == "LOGICAL_OR":
            acc = ((acc != 0) | (b != 0)).astype(t)
        elif name == "LOGICAL_XOR":
            acc = (acc != b).astype(t)
        else:
            ut = np.uint32 if np.dtype(t).itemsize == 4 else np.uint64
            x, y = acc.astype(ut), b.astype(ut)
            r = x & y if name == "BITWISE_AND" else x | y if name == "BITWISE_OR" else x ^ y
            acc = r.astype(t)
    return acc


def run_reduce(lib, op, ti, subtasks, n, device):
    t = TYPES[ti]
    out = np.zeros(n, dtype=t)
    rc = lib.rp_reduce(op, ti, subtasks, n, device, out.ctypes.data)
    return rc, out


@pytest.mark.parametrize("ti", range(6))
@pytest.mark.parametrize("op", range(10))
def test_builtin_reduction_cpu(lib, op, ti):
    n, subtasks = 1000, 5
    rc, got = run_reduce(lib, op, ti, subtasks, n, 0)
    assert rc == 0
    want = fold(op, TYPES[ti], [gen(op, TYPES[ti], s, n) for s in range(subtasks)])
    np.testing.assert_array_equal(got, want)


def test_reduction_planted_mistake_is_caught():
    # the fold above distinguishes LOGICAL_XOR (a != b) from a truth-value xor and MAX from MIN
    a = np.array([2, 0], dtype=np.int32)
    b = np.array([3, 0], dtype=np.int32)
    assert fold(OPS.index("LOGICAL_XOR"), np.int32, [a, b])[0] == 1
    assert fold(OPS.index("MAX"), np.int32, [a, b])[0] == 3 and fold(OPS.index("MIN"), np.int32, [a, b])[0] == 2
    # bitwise over floats acts on the integer values
    assert fold(OPS.index("BITWISE_OR"), np.float32, [np.float32([1.0]), np.float32([2.0])])[0] == 3.0


def chunks(sid, n, orders):
    lens, left = [], n
    for j in range(orders):
        l = left if j + 1 == orders else (int(_mix(sid, j)) % (left + 1) if left else 0)
        lens.append(l)
        left -= l
    return lens


def expected_redistribution(subtasks, n, orders):
    pieces = []
    for s in range(subtasks):
        off = 0
        for j, l in enumerate(chunks(s, n, orders)):
            pieces.append(((s + j) % orders, s, j, np.arange(s * n + off, s * n + off + l, dtype=np.uint32)))
            off += l
    pieces.sort(key=lambda p: (p[0], p[1], p[2]))
    body = np.concatenate([p[3] for p in pieces])
    meta = {}
    for o, _, _, v in pieces:
        meta[o] = meta.get(o, 0) + 4 * v.size
    return body, sorted((o, b) for o, b in meta.items() if b)


def run_redistribute(lib, subtasks, n, orders, device):
    out = np.zeros(subtasks * n, dtype=np.uint32)
    meta = np.zeros(2 * orders, dtype=np.uint32)
    cnt = C.c_uint32(0)
    rc = lib.rp_redistribute(subtasks, n, orders, device, out.ctypes.data, meta.ctypes.data, C.byref(cnt))
    return rc, out, [(int(meta[2 * k]), int(meta[2 * k + 1])) for k in range(cnt.value)]


@pytest.mark.parametrize("subtasks,n,orders", [(6, 500, 4), (3, 17, 5), (1, 64, 1)])
def test_redistribution_cpu(lib, subtasks, n, orders):
    rc, out, meta = run_redistribute(lib, subtasks, n, orders, 0)
    assert rc == 0
    body, want_meta = expected_redistribution(subtasks, n, orders)
    np.testing.assert_array_equal(out[:body.size], body)
    assert np.all(out[body.size:] == 0)
    assert meta == want_meta


@pytest.mark.parametrize("ti", range(6))
def test_sparse_reduction_cpu(lib, ti):
    """The sparse variant of the probe (about one element in 16 set per subtask), used by the GPU
    sentinel-compression tests, folds as numpy does on the CPU stubs."""
    n, subtasks = 1000, 4
    t = TYPES[ti]
    for op in range(10):
        out = np.zeros(n, dtype=t)
        rc = lib.rp_reduce_ex(op, ti, subtasks, n, 0, 1, out.ctypes.data)
        assert rc == 0, (op, ti)
        want = fold(op, t, [gen(op, t, s, n, sparse=True) for s in range(subtasks)])
        np.testing.assert_array_equal(out, want, err_msg=f"op {op} type {ti}")
    assert 0 < np.count_nonzero(gen(0, t, 0, n, sparse=True)) < n // 8
