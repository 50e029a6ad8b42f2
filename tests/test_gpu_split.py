"""GPU subtask splitting: imageFiltering and matrixMultiply tasks submitted with
canSplitGpuSubtasks run each subtask as pieces on the HIP stub, each piece its own kernel launch
(or work item in BATCHED mode) over its share of the rows, and the output stays bit-exact against
the oracle (reference pmSubtaskSplitter.cpp:118-140 for GPU split groups; the apps' row split,
imageFiltering.cpp:279-340 and matrixMultiplyBlas.cpp:67-88). The stub's subtask counter shows
that every piece ran: pieces x subtasks completions.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

MODES = [("plain", 0), ("resident", 1), ("batched", 2)]


@pytest.fixture
def split_gpu(gpu_runtime):
    yield gpu_runtime
    gpu_runtime.set_split()


@pytest.mark.parametrize("name,hip_mode", MODES)
@pytest.mark.parametrize("pieces", [3, 5])
def test_imagefilter_gpu_split_matches_oracle(split_gpu, name, hip_mode, pieces):
    U = split_gpu
    U.set_hip_mode(hip_mode)
    U.set_split(cpu=False, gpu=True, gpu_pieces=pieces)
    rng = np.random.default_rng(31 + pieces + 7 * hip_mode)
    H, W, r, tile = 300, 517, 2, 128
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
    f = U.ImageFilter(W, H, w, r, mode=0, shift=4, tile=tile)
    try:
        f.set_input(img)
        U.reset_stats()
        f.run(U.DEVICE_GPU)
        done = U.stats()["subtasks"]
        got = f.output()
    finally:
        f.close()
    subtasks = -(-H // tile) * -(-W // tile)
    assert done == pieces * subtasks, (done, pieces, subtasks)
    want = O.imagefilter(img, w, r, 0, 4)
    assert np.array_equal(got, want), f"{name}: {np.count_nonzero(got != want)} bytes differ"


@pytest.mark.parametrize("name,hip_mode", MODES)
def test_matmul_gpu_split_integer_exact(split_gpu, name, hip_mode):
    U = split_gpu
    U.set_hip_mode(hip_mode)
    U.set_split(cpu=False, gpu=True, gpu_pieces=3)
    rng = np.random.default_rng(41 + hip_mode)
    n, bs = 512, 128
    A = rng.integers(-8, 9, (n, n)).astype(np.float32)
    B = rng.integers(-8, 9, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    try:
        mm.set_inputs(A, B)
        U.reset_stats()
        mm.run(U.DEVICE_GPU)
        done = U.stats()["subtasks"]
        Cm = mm.output()
    finally:
        mm.close()
    assert done == 3 * (n // bs) ** 2
    assert np.array_equal(Cm, (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32))
