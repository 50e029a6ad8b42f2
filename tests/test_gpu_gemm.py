"""GPU GEMM (uh_gemm, split-BF16 3-term MFMA) and the matrixMultiply app against the FP64 oracle.

Tolerance: normwise relative error ||C - C64||_F / ||C64||_F <= 1e-5 (split-BF16 drops terms
below 2^-16 |a||b| per product; the expected error is ~1e-6). Small-integer matrices are exact in
split-BF16 and in fp32, so they must match the FP64 product bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _rel(c, ref):
    return float(np.linalg.norm(c.astype(np.float64) - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("M,N,K", [(128, 128, 128), (512, 512, 2048), (300, 517, 129), (1, 7, 3),
                                   (64, 1000, 40), (1000, 64, 33)])
@pytest.mark.parametrize("precision", [0, 2])
def test_gemm_shapes(gpu_runtime, M, N, K, precision):
    from unicorn_amd import hip as H
    rng = np.random.default_rng(M * 31 + N * 7 + K)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dA, dB, dC = H.DeviceBuffer.from_array(A), H.DeviceBuffer.from_array(B), H.DeviceBuffer(M * N * 4)
    H.gemm(dA, dB, dC, M, N, K, K, N, N, precision=precision)
    got = dC.to_array((M, N), np.float32)
    ref = O.matmul_f64(A, B)
    assert _rel(got, ref) <= 1e-5


def test_gemm_integer_exact(gpu_runtime):
    from unicorn_amd import hip as H
    rng = np.random.default_rng(3)
    M, N, K = 256, 384, 512
    A = rng.integers(-16, 17, (M, K)).astype(np.float32)
    B = rng.integers(-16, 17, (K, N)).astype(np.float32)
    dA, dB, dC = H.DeviceBuffer.from_array(A), H.DeviceBuffer.from_array(B), H.DeviceBuffer(M * N * 4)
    H.gemm(dA, dB, dC, M, N, K, K, N, N, precision=0)
    got = dC.to_array((M, N), np.float32)
    assert np.array_equal(got.astype(np.float64), O.matmul_f64(A, B))


def test_gemm_alpha_beta_strided(gpu_runtime):
    """C = -A B + C on sub-blocks of larger arrays (the LU trailing-update form)."""
    from unicorn_amd import hip as H
    rng = np.random.default_rng(4)
    ld = 700
    big_a = rng.uniform(-1, 1, (600, ld)).astype(np.float32)
    big_b = rng.uniform(-1, 1, (600, ld)).astype(np.float32)
    big_c = rng.uniform(-1, 1, (600, ld)).astype(np.float32)
    M, N, K = 512, 640, 500
    dA, dB, dC = H.DeviceBuffer.from_array(big_a), H.DeviceBuffer.from_array(big_b), H.DeviceBuffer.from_array(big_c)
    H.gemm(dA, dB, dC, M, N, K, ld, ld, ld, precision=0, alpha=-1.0, beta=1.0)
    got = dC.to_array((600, ld), np.float32)
    ref = big_c.astype(np.float64).copy()
    ref[:M, :N] -= O.matmul_f64(big_a[:M, :K], big_b[:K, :N])
    assert _rel(got[:M, :N], ref[:M, :N]) <= 1e-5
    assert np.array_equal(got[M:], big_c[M:]) and np.array_equal(got[:M, N:], big_c[:M, N:])


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_matmul_app(gpu_runtime, hip_mode):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    rng = np.random.default_rng(11)
    n, bs = 1024, 256
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    B = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    mm = U.MatMul(n, bs, 0)
    try:
        mm.set_inputs(A, B)
        mm.run(U.DEVICE_GPU)
        got = mm.output()
    finally:
        mm.close()
    assert _rel(got, O.matmul_f64(A, B)) <= 1e-5


@pytest.mark.parametrize("tile", range(8))
def test_gemm_every_tile_configuration(gpu_runtime, tile):
    """Each tile configuration (forced for the shape's class), plain launch and batched flush, on
    ragged shapes whose classes no other test uses."""
    import ctypes as C
    from unicorn_amd import hip as H
    L = H.lib()
    n_tiles = C.c_int(0)
    H.check(L.uh_gemm_tile_count(C.byref(n_tiles)), "uh_gemm_tile_count")
    assert n_tiles.value == 8
    for (M, N, K) in [(200, 700, 1100), (96, 80, 48)]:
        rng = np.random.default_rng(tile * 100 + M)
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        ref = O.matmul_f64(A, B)
        H.check(L.uh_gemm_set_tile(M, N, K, 0, tile), "uh_gemm_set_tile")
        dA, dB, dC = H.DeviceBuffer.from_array(A), H.DeviceBuffer.from_array(B), H.DeviceBuffer(M * N * 4)
        print(f"gemm tile {tile}: {M}x{N}x{K} plain", flush=True)
        H.gemm(dA, dB, dC, M, N, K, K, N, N, precision=0)
        assert _rel(dC.to_array((M, N), np.float32), ref) <= 1e-5
        s = H.Stream()
        bh = H.vp()
        H.check(L.uh_batch_create(C.byref(bh), s.h), "uh_batch_create")
        try:
            dC2 = H.DeviceBuffer(M * N * 4)
            print(f"gemm tile {tile}: {M}x{N}x{K} batched", flush=True)
            H.gemm(dA, dB, dC2, M, N, K, K, N, N, precision=0, stream=bh.value)
            H.check(L.uh_batch_flush(bh.value, 0), "uh_batch_flush")
            s.sync()
            assert np.array_equal(dC2.to_array((M, N), np.float32), dC.to_array((M, N), np.float32))
        finally:
            H.check(L.uh_batch_destroy(bh.value), "uh_batch_destroy")
