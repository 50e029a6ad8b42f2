"""GPU luDecomposition (panel, TRSM, split-BF16 / fp32 trailing update) and the refinement solve.

The float LU of the reference is not compared by the reference itself (luDecomposition.cpp:
593-596) and its cBLAS/cuBLAS rounding cannot be pinned, so parity is stated as tolerances:
  - factors vs the oracle's serialLUDecomposition restatement: max |dF| <= 1e-4 max |F|
    (fp32 blocked vs unblocked orderings; split-BF16 adds ~1e-6 relative per product);
  - backward error ||A - L U||_F / ||A||_F <= 1e-5 (fp32 level) with L, U from the factors;
  - refinement: normwise backward error <= 2^-24 after at most 5 steps, and forward error
    against a known FP64 solution <= 1e-6 (the FP64 oracle is exact: b = A x_true in FP64).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray((rng.random((n, n)) + n * np.eye(n)).astype(np.float32))


def _backward(A, F):
    n = A.shape[0]
    L = np.tril(F.astype(np.float64))
    Uu = np.triu(F.astype(np.float64), 1) + np.eye(n)
    return float(np.linalg.norm(A.astype(np.float64) - L @ Uu) / np.linalg.norm(A.astype(np.float64)))


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("precision", [0, 2])
def test_lu_factor(gpu_runtime, hip_mode, precision):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    n, B = 1024, 256
    A = _matrix(n, 7)
    lu = U.LU(n, B, precision)
    try:
        lu.set_matrix(A)
        lu.factor(U.DEVICE_GPU)
        F = lu.factors()
    finally:
        lu.close()
    Fo = O.lu(A)
    assert np.abs(F - Fo).max() <= 1e-4 * np.abs(Fo).max()
    assert _backward(A, F) <= 1e-5


def test_lu_panel_matches_oracle(gpu_runtime):
    """One diagonal block (S1 only) against the unblocked reference order."""
    from unicorn_amd import hip as H
    n = 512
    A = _matrix(n, 3)
    d = H.DeviceBuffer.from_array(A)
    H.check(H.lib().uh_lu_panel(d.ptr, n, n, None), "panel")
    F = d.to_array((n, n), np.float32, order="F")
    Fo = O.lu(A)
    assert np.abs(F - Fo).max() <= 1e-5 * np.abs(Fo).max()


@pytest.mark.parametrize("hip_mode", [1, 2])
def test_lu_solve_refinement(gpu_runtime, hip_mode):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    n, B = 2048, 512
    A = _matrix(n, 9)
    rng = np.random.default_rng(10)
    xt = rng.uniform(-1, 1, (n, 8))
    b = np.asfortranarray(A.astype(np.float64) @ xt)
    lu = U.LU(n, B, 0)
    try:
        lu.set_matrix(A)
        lu.factor(U.DEVICE_GPU)
        x, info = lu.solve(b, 5, U.DEVICE_GPU)
    finally:
        lu.close()
    assert info["steps"] <= 5
    assert info["backward_error"] <= 2.0 ** -24
    assert np.abs(x - xt).max() / np.abs(xt).max() <= 1e-6
    # the refined residual, recomputed on the host in FP64
    r = O.matvec_residual_f64(A, x, b)
    assert np.abs(r).max() / (info["a_inf"] * np.abs(x).max() + np.abs(b).max()) <= 2.0 ** -24
