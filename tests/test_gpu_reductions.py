"""Built-in reductions and data redistribution through the HIP stub.

The probe task of tests/test_reductions.py runs with its kernel-kind GPU callback: in PLAIN mode
the subtasks' outputs come back to host shadows and fold there; in RESIDENT and BATCHED mode the
shadows stay in HBM and fold with uh_reduce, for all ten operations over all six element types,
equal to numpy's fold. Redistribution: the GPU callback writes its chunk lengths into a
SUBTASK_TO_POST_SUBTASK scratch buffer that the runtime copies back for the dataRedistribution
callback; the redistributed space must equal the concatenation by (order, subtask).
"""
import numpy as np
import pytest

from test_reductions import TYPES, expected_redistribution, fold, gen, load, run_reduce, run_redistribute

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_runtime):
    return load(gpu_runtime)


@pytest.mark.parametrize("hip_mode", [0, 1, 2])
@pytest.mark.parametrize("ti", range(6))
def test_builtin_reductions_on_gpu(gpu_runtime, lib, hip_mode, ti):
    gpu_runtime.set_hip_mode(hip_mode)
    n, subtasks = 3000, 6
    for op in range(10):
        rc, got = run_reduce(lib, op, ti, subtasks, n, 1)
        assert rc == 0, (op, ti)
        want = fold(op, TYPES[ti], [gen(op, TYPES[ti], s, n) for s in range(subtasks)])
        np.testing.assert_array_equal(got, want, err_msg=f"op {op} type {ti} mode {hip_mode}")


@pytest.mark.parametrize("device", [1, 2])
@pytest.mark.parametrize("hip_mode", [0, 1, 2])
def test_redistribution_on_gpu(gpu_runtime, lib, device, hip_mode):
    """PLAIN stages the subtasks through host memory and redistributes there; a GPU-only task in
    the resident modes keeps its output in HBM and is redistributed there (one copy kernel)."""
    gpu_runtime.set_hip_mode(hip_mode)
    subtasks, n, orders = 8, 1000, 5
    rc, out, meta = run_redistribute(lib, subtasks, n, orders, device)
    assert rc == 0
    body, want_meta = expected_redistribution(subtasks, n, orders)
    np.testing.assert_array_equal(out[:body.size], body)
    assert np.all(out[body.size:] == 0)
    assert meta == want_meta
