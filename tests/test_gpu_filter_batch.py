"""imageFiltering work items of different kinds in one batch (uh_imagefilter recorded into a batch
handle, one uh_batch_flush): radius 1 and 2 in both modes and a radius-3 item, so the persistent
launch runs the kernel variant that dispatches per item (both engines' tile paths and the direct
path in one grid). Every item's output is bit-exact against the oracle, with each engine.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


class FilterArgs(C.Structure):
    _fields_ = [("in_", C.c_void_p), ("in_stride", C.c_uint64), ("out", C.c_void_p), ("out_stride", C.c_uint64),
                ("row0", C.c_int32), ("col0", C.c_int32), ("rows", C.c_int32), ("cols", C.c_int32),
                ("view_last_row", C.c_int32), ("view_first_col", C.c_int32),
                ("height", C.c_int32), ("width", C.c_int32),
                ("radius", C.c_int32), ("mode", C.c_int32), ("shift", C.c_int32), ("pad", C.c_int32),
                ("weights", C.c_int8 * 1024)]


def _args(d_in, d_out, H, W, r, mode, shift, w):
    """The whole image as one view, as the app's GPU callback describes a one-subtask task."""
    a = FilterArgs()
    a.in_, a.in_stride, a.out, a.out_stride = d_in.ptr, W * 3, d_out.ptr, W * 3
    a.row0, a.col0, a.rows, a.cols = 0, 0, H, W
    a.view_last_row, a.view_first_col = H - 1, 0
    a.height, a.width = H, W
    a.radius, a.mode, a.shift = r, mode, shift
    for i, v in enumerate(w.ravel()):
        a.weights[i] = int(v)
    return a


@pytest.mark.parametrize("engine", [1, 0])
def test_mixed_filter_items_in_one_batch(gpu_runtime, engine):
    from unicorn_amd import hip as H
    L = H.lib()
    L.uh_imagefilter.argtypes = [C.c_void_p, C.c_void_p]
    L.uh_imagefilter.restype = C.c_int
    prev = C.c_int(0)
    H.check(L.uh_imagefilter_engine(engine, C.byref(prev)), "engine")
    rng = np.random.default_rng(11 + engine)
    Hh, W = 300, 517
    img = rng.integers(0, 256, (Hh, W, 3), dtype=np.uint8)
    kinds = [(1, 0), (2, 1), (1, 1), (3, 0), (2, 0)]
    d_in = H.DeviceBuffer.from_array(img)
    outs, wants, keep = [], [], []
    s = H.Stream()
    bh = H.vp()
    H.check(L.uh_batch_create(C.byref(bh), s.h), "uh_batch_create")
    try:
        for r, mode in kinds:
            d = 2 * r + 1
            w = (rng.integers(-4, 5, d * d) if mode else rng.integers(-128, 128, d * d)).astype(np.int8)
            d_out = H.DeviceBuffer(Hh * W * 3)
            a = _args(d_in, d_out, Hh, W, r, mode, 4, w)
            keep.append(a)
            print(f"filter batch: engine {engine} radius {r} mode {mode}", flush=True)
            H.check(L.uh_imagefilter(C.byref(a), bh.value), "uh_imagefilter")
            outs.append(d_out)
            wants.append(O.imagefilter(img, w, r, mode, 4))
        H.check(L.uh_batch_flush(bh.value, 0), "uh_batch_flush")
        s.sync()
        for (r, mode), d_out, want in zip(kinds, outs, wants):
            got = d_out.to_array((Hh, W, 3), np.uint8)
            assert np.array_equal(got, want), f"radius {r} mode {mode}: {np.count_nonzero(got != want)} bytes differ"
    finally:
        H.check(L.uh_batch_destroy(bh.value), "uh_batch_destroy")
        H.check(L.uh_imagefilter_engine(prev.value, None), "engine")
        for d_out in outs:
            d_out.free()
        d_in.free()
