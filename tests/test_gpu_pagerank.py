"""GPU pageRank (pull SpMV over the HBM-resident in-link CSR) against the oracle.

Tolerance: relative L1 <= 1e-5 against the oracle's restatement of the reference (push per
subtask + REDUCE_ADD in subtask order); the error against the FP64 oracle is asserted too.
Closed forms pin the recurrence: a ring gives 0.85 everywhere, a star gives the hub
0.85 * leaves after the first iteration, and a graph of dangling pages gives zero.
All three graph layouts run: binned (the default: propagation-blocked link values with
fixed-point row sums), source-blocked, and the plain in-link CSR. In the blocked layout the GPU
kernels and the CPU callback add in the same order, and in the binned layout both add fixed-point
integers, so in both their ranks are bitwise equal.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


LAYOUTS = ["binned", "blocked", "csr"]


@pytest.fixture
def layout(gpu_runtime, request):
    gpu_runtime.pr_set_layout(request.param)
    yield request.param
    gpu_runtime.pr_set_layout("binned")


@pytest.mark.parametrize("layout", LAYOUTS, indirect=True)
@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_pagerank_matches_oracle(gpu_runtime, hip_mode, layout):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rp, col = U.webgen(200_000, 11)
    pr = U.PageRank(rp, col, 50_000)
    try:
        assert pr.layout == layout
        print(f"pagerank: layout {layout}, mode {hip_mode}", flush=True)
        pr.run(20, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 20, 50_000)) <= 1e-5
    assert _rel_l1(y, O.pagerank_f64(rp, col, 20)) <= 1e-5


@pytest.mark.parametrize("layout", LAYOUTS, indirect=True)
def test_pagerank_compact_views(gpu_runtime, layout):
    U = gpu_runtime
    U.set_hip_mode(1, compact=True)
    rp, col = U.webgen(30_000, 5)
    pr = U.PageRank(rp, col, 7_000)      # ragged last subtask; subtasks share row groups
    try:
        assert pr.layout == layout
        pr.run(4, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
        U.set_hip_mode(2)
    assert _rel_l1(y, O.pagerank(rp, col, 4, 7_000)) <= 1e-5


def _csr(n, edges):
    rp = np.zeros(n + 1, dtype=np.uint32)
    for s, _ in edges:
        rp[s + 1] += 1
    rp = np.cumsum(rp).astype(np.uint32)
    col = np.zeros(max(1, len(edges)), dtype=np.uint32)
    pos = rp[:-1].copy()
    for s, t in sorted(edges):
        col[pos[s]] = t
        pos[s] += 1
    return rp, col[:len(edges)]


@pytest.mark.parametrize("layout", LAYOUTS, indirect=True)
def test_pagerank_closed_forms(gpu_runtime, layout):
    U = gpu_runtime
    U.set_hip_mode(2)
    n = 1000
    rp, col = _csr(n, [(i, (i + 1) % n) for i in range(n)])       # ring
    pr = U.PageRank(rp, col, 300)
    pr.run(3, U.DEVICE_GPU)
    np.testing.assert_allclose(pr.ranks(), np.float32(0.85) ** 3, rtol=1e-6)   # x_k = 0.85^k
    pr.close()
    rp, col = _csr(n, [(i, 0) for i in range(1, n)])              # star into page 0
    pr = U.PageRank(rp, col, 256)
    pr.run(1, U.DEVICE_GPU)
    y = pr.ranks()
    pr.close()
    assert abs(y[0] - 0.85 * (n - 1)) <= 1e-5 * 0.85 * (n - 1)
    assert np.all(y[1:] == 0)
    rp = np.zeros(n + 1, dtype=np.uint32)                           # all dangling
    pr = U.PageRank(rp, np.zeros(0, dtype=np.uint32), 128)
    pr.run(2, U.DEVICE_GPU)
    assert np.all(pr.ranks() == 0)
    pr.close()


def test_pagerank_cpu_gpu_mix(gpu_runtime):
    U = gpu_runtime
    U.set_hip_mode(1)
    rp, col = U.webgen(50_000, 2)
    pr = U.PageRank(rp, col, 5_000)
    try:
        pr.run(5, U.DEVICE_BOTH)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 5, 5_000)) <= 1e-5


@pytest.mark.parametrize("layout", ["binned", "blocked"], indirect=True)
@pytest.mark.parametrize("hip_mode", [0, 2])
def test_pagerank_cpu_gpu_bitwise(gpu_runtime, hip_mode, layout):
    """The layouts whose CPU callback adds in the kernels' order (blocked) or in an order-free
    fixed point (binned): CPU and GPU ranks are the same bits."""
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rp, col = U.webgen(100_000, 4)
    ys = []
    for device in (U.DEVICE_CPU, U.DEVICE_GPU):
        pr = U.PageRank(rp, col, 30_000)
        try:
            assert pr.layout == layout
            pr.run(3, device)
            ys.append(pr.ranks())
        finally:
            pr.close()
    assert np.array_equal(ys[0], ys[1])
    assert _rel_l1(ys[1], O.pagerank(rp, col, 3, 30_000)) <= 1e-5


@pytest.mark.parametrize("layout", ["blocked"], indirect=True)
def test_pagerank_csr_fallback_for_heavy_rows(gpu_runtime, layout):
    U = gpu_runtime
    U.set_hip_mode(2)
    n = 20_000
    rp0, col0 = U.webgen(n, 9)
    src = np.repeat(np.arange(n), np.diff(rp0.astype(np.int64)))
    edges = set(zip(src.tolist(), col0.tolist()))
    edges |= {(s, 12_345) for s in range(400)}   # 400 in-links from source block 0: > 255
    rp, col = _csr(n, sorted(edges))
    pr = U.PageRank(rp, col, 3000)
    try:
        assert not pr.blocked
        pr.run(2, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 2, 3000)) <= 1e-5


@pytest.mark.parametrize("hip_mode", [0, 2])
def test_pagerank_from_webdump(gpu_runtime, hip_mode, tmp_path):
    """A graph written in the reference's on-disk web dump format and loaded back (pmMapFile'd
    records → resident graph) gives the oracle's ranks."""
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rp, col = U.webgen(150_000, 12)
    U.webdump_write(tmp_path, rp, col, 100, 40_000)
    pr = U.PageRank.from_webdump(tmp_path, 50_000)
    try:
        assert pr.layout == "binned"
        pr.run(10, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 10, 50_000)) <= 1e-5


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_pagerank_push_reference_decomposition(gpu_runtime, hip_mode):
    """The reference's decomposition on the GPU: source-page subtasks push (float atomics) into
    private full-length vectors; REDUCE_ADD folds them in HBM (resident modes) or on the host
    after their D2H (PLAIN)."""
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rp, col = U.webgen(120_000, 14)
    pr = U.PageRankPush(rp, col, 25_000)
    try:
        pr.run(8, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 8, 25_000)) <= 1e-5


def test_pagerank_push_from_webdump(gpu_runtime, tmp_path):
    U = gpu_runtime
    U.set_hip_mode(2)
    rp, col = U.webgen(80_000, 15)
    U.webdump_write(tmp_path, rp, col, 100, 20_000)
    pr = U.PageRankPush.from_webdump(tmp_path, 20_000)
    try:
        pr.run(6, U.DEVICE_GPU)
        y = pr.ranks()
    finally:
        pr.close()
    assert _rel_l1(y, O.pagerank(rp, col, 6, 20_000)) <= 1e-5
