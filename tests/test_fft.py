"""fft on the CPU: the oracle's 2-D DFT pinned by independent checks, and the app's CPU path
(row task, then column task over scattered subscriptions, with CPU subtask splitting) against it.

Pins of oracle.fft2 (fftSerial, reference fft.cpp:115-155; FFTW's forward transform restated by
its definition):
  - brute force: the DFT matrices exp(-2 pi i j k / n) applied to rows and columns (tiny sizes);
  - numpy.fft.fft2 in complex128 (an independent implementation);
  - closed forms: a unit impulse at (r0, c0) gives W_R^(r0 u) W_C^(c0 v); a constant gives R C at
    (0, 0) only; exp(+2 pi i (u0 r / R + v0 c / C)) gives R C at (u0, v0) only; the reference's
    own input x = (i, i + 1), i = r C + c (fft.cpp:612-630) gives
    X = (1 + i) F + i R C d(u) d(v) with F[u][v] = C^2 S_R(u) d(v) + R d(u) S_C(v), where
    S_N(0) = N (N - 1) / 2 and S_N(k) = N / (exp(-2 pi i k / N) - 1) otherwise.
A planted sign error (the inverse transform) fails the brute-force and closed-form checks.
Tolerance of the float app path: ||X - X64||_F / ||X64||_F <= 1e-5 (the float-output bound).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

TOL = 1e-5


def _brute(x):
    R, Cc = x.shape
    wr = np.exp(-2j * np.pi * np.outer(np.arange(R), np.arange(R)) / R)
    wc = np.exp(-2j * np.pi * np.outer(np.arange(Cc), np.arange(Cc)) / Cc)
    return wr @ x.astype(np.complex128) @ wc


def _rand(R, Cc, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, (R, Cc)) + 1j * rng.uniform(-1, 1, (R, Cc))).astype(np.complex64)


def _ramp(R, Cc):
    i = np.arange(R * Cc, dtype=np.float64).reshape(R, Cc)
    return (i + 1j * (i + 1)).astype(np.complex64)


def _S(N):
    k = np.arange(N)
    s = np.empty(N, np.complex128)
    s[0] = N * (N - 1) / 2
    s[1:] = N / (np.exp(-2j * np.pi * k[1:] / N) - 1)
    return s


def ramp_closed_form(R, Cc):
    dR = (np.arange(R) == 0).astype(np.float64)
    dC = (np.arange(Cc) == 0).astype(np.float64)
    F = Cc * Cc * np.outer(_S(R), dC) + R * np.outer(dR, _S(Cc))
    return (1 + 1j) * F + 1j * R * Cc * np.outer(dR, dC)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", [(1, 1), (2, 4), (8, 8), (4, 16), (32, 16), (16, 64)])
def test_oracle_brute_force(shape):
    x = _rand(*shape, 1)
    assert _rel(O.fft2(x), _brute(x)) < 1e-13


def test_oracle_numpy():
    x = _rand(256, 512, 2)
    assert _rel(O.fft2(x), np.fft.fft2(x.astype(np.complex128))) < 1e-14


def test_oracle_closed_forms():
    R, Cc = 32, 64
    u, v = np.arange(R)[:, None], np.arange(Cc)[None, :]
    for r0, c0 in ((0, 0), (3, 5), (31, 63)):
        x = np.zeros((R, Cc), np.complex64)
        x[r0, c0] = 1
        want = np.exp(-2j * np.pi * (r0 * u / R + c0 * v / Cc))
        assert np.abs(O.fft2(x) - want).max() < 1e-13
    got = O.fft2(np.full((R, Cc), 2.5 - 1j, np.complex64))
    want = np.zeros((R, Cc), np.complex128)
    want[0, 0] = R * Cc * (2.5 - 1j)
    assert np.abs(got - want).max() < 1e-11
    u0, v0 = 7, 40
    x = np.exp(2j * np.pi * (u0 * np.arange(R)[:, None] / R + v0 * np.arange(Cc)[None, :] / Cc)).astype(np.complex64)
    got = O.fft2(x)
    assert abs(got[u0, v0] - R * Cc) < 1e-3
    got[u0, v0] = 0
    assert np.abs(got).max() < 1e-3


@pytest.mark.parametrize("shape", [(8, 8), (64, 32), (512, 512)])
def test_oracle_reference_input_closed_form(shape):
    R, Cc = shape
    want = ramp_closed_form(R, Cc)
    assert _rel(O.fft2(_ramp(R, Cc)), want) < 1e-12
    assert _rel(_brute(_ramp(8, 8)), ramp_closed_form(8, 8)) < 1e-12


def test_planted_sign_error_is_caught():
    """The inverse transform (a sign slip in the exponent) disagrees with every pin above."""
    x = _rand(16, 32, 3)
    wrong = np.conj(O.fft2(np.conj(x)))        # sum x exp(+2 pi i ...)
    assert _rel(wrong, _brute(x)) > 0.5
    assert _rel(np.conj(O.fft2(np.conj(_ramp(16, 32)))), ramp_closed_form(16, 32)) > 0.5


@pytest.fixture(params=[0, 3, -1], ids=["split-per-stub", "split-3", "unsplit"])
def cpu_runtime(runtime, request):
    """CPU subtasks split into one piece per CPU stub (the reference's split group), into 3
    uneven pieces (R / 3 each, the last one the rest), or not split."""
    p = request.param
    runtime.set_split(cpu=p >= 0, cpu_pieces=max(p, 0))
    yield runtime
    runtime.set_split()


def _run(U, x, R, device):
    f = U.Fft2D(x.shape[0], x.shape[1], R)
    try:
        f.set_input(x)
        f.run(device)
        return f.output()
    finally:
        f.close()


@pytest.mark.parametrize("shape,R", [((64, 64), 16), ((128, 256), 32), ((256, 64), 64), ((512, 512), 512)])
def test_app_cpu_matches_oracle(cpu_runtime, shape, R):
    x = _rand(*shape, 5)
    got = _run(cpu_runtime, x, R, cpu_runtime.DEVICE_CPU)
    assert got.dtype == np.complex64
    assert _rel(got, O.fft2(x)) <= TOL


def test_app_cpu_reference_input(cpu_runtime):
    R = Cc = 1024
    got = _run(cpu_runtime, _ramp(R, Cc), 512, cpu_runtime.DEVICE_CPU)
    assert _rel(got, ramp_closed_form(R, Cc)) <= TOL


def test_app_rejects_bad_shapes(runtime):
    for args in ((96, 64, 16), (64, 64, 128), (64, 64, 24), (16384, 16, 16)):
        with pytest.raises(RuntimeError):
            runtime.Fft2D(*args)
