"""The kernel-kind GPU callback contract (subtask_gpu_hip), in all three HIP stub modes.

tests/native/kernel_kind_probe.hip registers a task whose GPU callback is a __global__ kernel.
Its output encodes what the kernel saw: the device copy of taskConf (n, magic, threads), the
device copy of pmDeviceInfo (deviceType == GPU_HIP), the launch geometry set with
pmSetGpuLaunchConf (blockDim.x), the PRE_SUBTASK_TO_SUBTASK scratch buffer filled on the host by
the data-distribution callback (read through the device-side pmGetScratchBuffer), the memory of
pmReserveGpuGlobalMem (written and read back), and the subtask's output slice. The CPU callback
takes the same scratch buffer inside subtask_cpu. Reference contract: pmCudaInterface.cu:532-543,
pmExecutionStub.cpp:3300-3368, pmDispatcherGPU.cpp:161-205.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

PROBE = Path(__file__).resolve().parent / "native" / "libkernel_kind_probe.so"


def _lib():
    if not PROBE.exists():
        pytest.fail(f"{PROBE} is not built (unicorn_amd.build.build_test_support)")
    L = C.CDLL(str(PROBE))
    L.kk_probe_run.restype = C.c_int
    L.kk_probe_run.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
    return L


def _expected(n, subtasks, magic, threads, gpu):
    s = np.arange(subtasks, dtype=np.uint64)[:, None]
    i = np.arange(n, dtype=np.uint64)[None, :]
    sc = (s * 1000003 + i) & 0xFFFFFFFF
    v = (sc ^ magic) + threads * 16 + subtasks + (1 if gpu else 0)
    return (v & 0xFFFFFFFF).astype(np.uint32).reshape(-1)


def _run(n, subtasks, magic, threads, device):
    out = np.zeros(n * subtasks, dtype=np.uint32)
    rc = _lib().kk_probe_run(n, subtasks, magic, threads, device, out.ctypes.data)
    assert rc == 0, f"kk_probe_run: status {rc}"
    return out


def test_kernel_kind_probe_cpu(runtime):
    got = _run(1000, 7, 0x5A5A1234, 64, 0)
    assert np.array_equal(got, _expected(1000, 7, 0x5A5A1234, 64, False))


@pytest.mark.gpu
@pytest.mark.parametrize("hip_mode", [0, 1, 2], ids=["plain", "resident", "batched"])
@pytest.mark.parametrize("n,subtasks,threads", [(1000, 7, 64), (4096, 33, 256), (1, 1, 1), (70_001, 3, 1024)])
def test_kernel_kind_probe_gpu(gpu_runtime, hip_mode, n, subtasks, threads):
    U = gpu_runtime
    U.set_hip_mode(hip_mode)
    try:
        for magic in (0x5A5A1234, 0x0F0F0F0F):       # two tasks in a row: each has its own taskConf copy
            print(f"kernel-kind probe: mode {hip_mode}, n {n}, subtasks {subtasks}, threads {threads}", flush=True)
            got = _run(n, subtasks, magic, threads, 1)
            want = _expected(n, subtasks, magic, threads, True)
            assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} of {got.size} differ"
    finally:
        U.set_hip_mode(U.PM_HIP_MODE_BATCHED)
