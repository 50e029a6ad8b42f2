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


def _unit_lower_plus_diag(n, seed):
    """Lower triangle with a dominant diagonal (the L of the packed factors)."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.uniform(-1, 1, (n, n))) / n
    L[np.diag_indices(n)] = rng.uniform(1, 2, n)
    return L


def _padded(A, ld, fill):
    """Column-major copy of A with leading dimension ld; rows past A's are `fill`."""
    out = np.full((ld, A.shape[1]), fill, np.float32, order="F")
    out[: A.shape[0]] = A
    return out


@pytest.mark.parametrize("n,m,pad", [(512, 512, 0), (200, 77, 3), (33, 5, 1), (31, 40, 0)])
def test_trsm_lower_left(gpu_runtime, n, m, pad):
    """X <- L^-1 X (the first half of stage 2). The strictly upper triangle is NaN: the kernel
    must read only the lower triangle, as the packed factors keep U there."""
    from unicorn_amd import hip as H
    L = _unit_lower_plus_diag(n, 11)
    T = L.astype(np.float32)
    T[np.triu_indices(n, 1)] = np.nan
    X = np.random.default_rng(12).uniform(-1, 1, (n, m)).astype(np.float32)
    dT = H.DeviceBuffer.from_array(_padded(T, n + pad, np.nan))
    dX = H.DeviceBuffer.from_array(_padded(X, n + 2 * pad, 0.0))
    H.check(H.lib().uh_trsm_lln(dT.ptr, n + pad, dX.ptr, n + 2 * pad, n, m, None), "trsm_lln")
    got = dX.to_array((n + 2 * pad, m), np.float32, order="F")[:n]
    want = np.linalg.solve(np.tril(L.astype(np.float32).astype(np.float64)), X.astype(np.float64))
    assert np.isfinite(got).all()
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("m,n,pad", [(512, 512, 0), (77, 200, 3), (5, 33, 1), (40, 31, 0)])
def test_trsm_right_upper_unit(gpu_runtime, m, n, pad):
    """X <- X U^-1 with U unit upper (the second half of stage 2). The diagonal and the lower
    triangle are NaN: the kernel must read only the strictly upper triangle."""
    from unicorn_amd import hip as H
    rng = np.random.default_rng(13)
    U = np.triu(rng.uniform(-1, 1, (n, n)), 1) / n + np.eye(n)
    T = U.astype(np.float32)
    T[np.tril_indices(n)] = np.nan
    X = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dT = H.DeviceBuffer.from_array(_padded(T, n + pad, np.nan))
    dX = H.DeviceBuffer.from_array(_padded(X, m + 2 * pad, 0.0))
    H.check(H.lib().uh_trsm_run(dT.ptr, n + pad, dX.ptr, m + 2 * pad, m, n, None), "trsm_run")
    got = dX.to_array((m + 2 * pad, n), np.float32, order="F")[:m]
    Uf = np.triu(U.astype(np.float32).astype(np.float64), 1) + np.eye(n)
    want = np.linalg.solve(Uf.T, X.astype(np.float64).T).T
    assert np.isfinite(got).all()
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("n", [45, 300])
def test_lu_panel_ragged(gpu_runtime, n):
    """Blocks whose order is not a multiple of the 32-column sub-panel, with lda > n."""
    from unicorn_amd import hip as H
    A = _matrix(n, 5)
    lda = n + 7
    d = H.DeviceBuffer.from_array(_padded(A, lda, 0.0))
    H.check(H.lib().uh_lu_panel(d.ptr, lda, n, None), "panel")
    F = d.to_array((lda, n), np.float32, order="F")[:n]
    Fo = O.lu(A)
    assert np.abs(F - Fo).max() <= 1e-5 * np.abs(Fo).max()


@pytest.mark.parametrize("n,nrhs", [(700, 3), (512, 8), (64, 1)])
def test_lu_solve_f32(gpu_runtime, n, nrhs):
    """L U z = r with the packed factors (blocks of 512 rows, ragged last block)."""
    from unicorn_amd import hip as H
    L = _unit_lower_plus_diag(n, 21).astype(np.float32)
    rng = np.random.default_rng(22)
    U = (np.triu(rng.uniform(-1, 1, (n, n)), 1) / n).astype(np.float32)
    F = np.asfortranarray(L + U)
    r = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)).astype(np.float32))
    dF = H.DeviceBuffer.from_array(F)
    dr = H.DeviceBuffer.from_array(r)
    H.check(H.lib().uh_lu_solve_f32(dF.ptr, n, n, dr.ptr, nrhs, None), "lu_solve_f32")
    got = dr.to_array((n, nrhs), np.float32, order="F")
    L64 = L.astype(np.float64)
    U64 = U.astype(np.float64) + np.eye(n)
    want = np.linalg.solve(U64, np.linalg.solve(L64, r.astype(np.float64)))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
