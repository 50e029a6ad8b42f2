"""GPU parity of the imageFiltering device subtask path against the oracle (bit-exact).

Every HIP stub mode runs the same app callbacks through the C-ABI driver
(libunicorn_apps.so -> libpmlib.so -> libunicorn_hip.so); the output must equal the oracle's
restatement of the reference CPU callback byte for byte (reference imageFiltering.cpp:379-472).
"""
import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

MODES = [
    ("plain", 0, False, False),
    ("resident", 1, False, False),
    ("resident_compact", 1, True, False),
    ("batched", 2, False, False),
    ("batched_graph", 2, False, True),
]


def _case(rng, H, W, r, mode):
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    if mode == 0:
        w = rng.integers(-128, 128, (2 * r + 1) ** 2).astype(np.int8)
    else:
        w = rng.integers(-4, 5, (2 * r + 1) ** 2).astype(np.int8)
    return img, w


@pytest.fixture(params=[1, 0], ids=["mfma", "dot4"])
def engine(gpu_runtime, request):
    """The kernel engine of radius 1 and 2 (uh_imagefilter_engine): int8 MFMA or v_dot4."""
    from unicorn_amd import hip as H
    import ctypes as C
    prev = C.c_int()
    H.check(H.lib().uh_imagefilter_engine(request.param, C.byref(prev)), "engine")
    yield request.param
    H.check(H.lib().uh_imagefilter_engine(prev.value, None), "engine")


@pytest.mark.parametrize("name,hip_mode,compact,graphs", MODES)
@pytest.mark.parametrize("H,W,r,mode,tile", [
    (300, 517, 2, 0, 128),     # ragged tiles, tile not a multiple of the kernel tile
    (300, 517, 2, 1, 128),
    (64, 64, 1, 0, 2048),      # one subtask, radius 1
    (4, 3, 2, 1, 2),           # image smaller than the filter window
    (257, 1030, 3, 0, 256),    # radius 3: direct kernel
    (1030, 2100, 2, 1, 512),
    (333, 1001, 1, 1, 300),    # radius 1, odd widths: staged rows at every byte alignment
    (200, 1000, 2, 1, 200),    # 3000-byte rows: subtask columns alternate 16- and 8-byte aligned
])
def test_imagefilter_gpu_matches_oracle(gpu_runtime, engine, name, hip_mode, compact, graphs, H, W, r, mode, tile):
    U = gpu_runtime
    U.set_hip_mode(hip_mode, compact=compact, graphs=graphs)
    rng = np.random.default_rng(H * 7 + W * 13 + r + 100 * mode + tile)
    img, w = _case(rng, H, W, r, mode)
    f = U.ImageFilter(W, H, w, r, mode=mode, shift=4, tile=tile)
    try:
        f.set_input(img)
        f.run(U.DEVICE_GPU)
        got = f.output()
    finally:
        f.close()
    want = O.imagefilter(img, w, r, mode, 4)
    assert np.array_equal(got, want), f"{name}: {np.count_nonzero(got != want)} bytes differ"


def test_imagefilter_gpu_repeat_resident(gpu_runtime, engine):
    """The same task twice on a resident input (second run reuses the HBM mirror)."""
    U = gpu_runtime
    U.set_hip_mode(2, graphs=True)
    rng = np.random.default_rng(5)
    img, w = _case(rng, 2048, 2048, 2, 1)
    f = U.ImageFilter(2048, 2048, w, 2, mode=1, shift=4, tile=1024)
    try:
        f.set_input(img)
        U.reset_stats()
        f.run(U.DEVICE_GPU)
        f.run(U.DEVICE_GPU)
        s = U.stats()
        got = f.output()
    finally:
        f.close()
    assert np.array_equal(got, O.imagefilter(img, w, 2, 1, 4))
    # the input was uploaded once; nothing is staged per subtask
    assert s["h2d_bytes"] <= 2 * 2048 * 2048 * 3 * 2


def test_imagefilter_gpu_cpu_gpu_mix(gpu_runtime, engine):
    """CPU and GPU stubs on one task (the reference's CPU+GPU mode)."""
    U = gpu_runtime
    U.set_hip_mode(1)
    rng = np.random.default_rng(9)
    img, w = _case(rng, 700, 900, 2, 0)
    f = U.ImageFilter(900, 700, w, 2, mode=0, tile=128)
    try:
        f.set_input(img)
        f.run(U.DEVICE_BOTH)
        got = f.output()
    finally:
        f.close()
    assert np.array_equal(got, O.imagefilter(img, w, 2, 0, 4))
