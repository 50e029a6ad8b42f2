"""ctypes bindings of the HIP C-ABI (libunicorn_hip.so) for tests and measurement.

Device memory, streams and events are the library's own (hipMalloc/hipStream behind uh_*), so a
test can call a kernel launcher exactly as the pm:: runtime's GPU callbacks do.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip_lib

vp, i32, u32, u64, f32 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_float


class uh_gemm_args(C.Structure):
    _fields_ = [("A", vp), ("B", vp), ("C", vp), ("lda", u64), ("ldb", u64), ("ldc", u64),
                ("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32), ("precision", C.c_int32),
                ("alpha", f32), ("beta", f32)]


class uh_pagerank_args(C.Structure):
    _fields_ = [("row_ptr", vp), ("col", vp), ("deg", vp), ("w", vp), ("y", vp), ("w_next", vp),
                ("row0", u32), ("rows", u32), ("damping", f32), ("pad", u32),
                ("bcnt", vp), ("bseg", vp)]


_declared = False


def lib() -> C.CDLL:
    global _declared
    L = hip_lib()
    if not _declared:
        sig = {
            "uh_last_error": (C.c_char_p, []),
            "uh_device_count": (i32, [C.POINTER(i32)]),
            "uh_malloc": (i32, [C.POINTER(vp), C.c_size_t]),
            "uh_free": (i32, [vp]),
            "uh_memcpy_h2d": (i32, [vp, vp, C.c_size_t, vp]),
            "uh_memcpy_d2h": (i32, [vp, vp, C.c_size_t, vp]),
            "uh_memcpy_d2d": (i32, [vp, vp, C.c_size_t, vp]),
            "uh_fill_u32": (i32, [vp, u32, C.c_size_t, vp]),
            "uh_stream_create": (i32, [C.POINTER(vp)]),
            "uh_stream_destroy": (i32, [vp]),
            "uh_stream_sync": (i32, [vp]),
            "uh_device_sync": (i32, []),
            "uh_event_create": (i32, [C.POINTER(vp)]),
            "uh_event_destroy": (i32, [vp]),
            "uh_event_record": (i32, [vp, vp]),
            "uh_event_sync": (i32, [vp]),
            "uh_event_elapsed_ms": (i32, [C.POINTER(f32), vp, vp]),
            "uh_gemm": (i32, [C.POINTER(uh_gemm_args), vp]),
            "uh_gemm_set_tile": (i32, [i32, i32, i32, i32, i32]),
            "uh_gemm_tile_count": (i32, [C.POINTER(i32)]),
            "uh_pagerank_pull": (i32, [C.POINTER(uh_pagerank_args), vp]),
            "uh_pagerank_pull_noalloc": (i32, [C.POINTER(uh_pagerank_args), vp]),
            "uh_pagerank_init_w": (i32, [vp, vp, u32, f32, f32, vp]),
            "uh_lu_panel": (i32, [vp, u64, i32, vp]),
            "uh_trsm_lln": (i32, [vp, u64, vp, u64, i32, i32, vp]),
            "uh_trsm_run": (i32, [vp, u64, vp, u64, i32, i32, vp]),
            "uh_residual_f64": (i32, [vp, u64, i32, vp, vp, i32, vp, vp, vp]),
            "uh_lu_solve_f32": (i32, [vp, u64, i32, vp, i32, vp]),
            "uh_batch_create": (i32, [C.POINTER(vp), vp]),
            "uh_batch_destroy": (i32, [vp]),
            "uh_batch_flush": (i32, [vp, i32]),
            "uh_launch_counters": (i32, [C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
            "uh_imagefilter_engine": (i32, [i32, C.POINTER(i32)]),
            "uh_transpose32": (i32, [vp, u64, vp, u64, u32, u32, vp]),
            "uh_scan_i32": (i32, [vp, vp, u64, vp]),
            "uh_spmm_coo_f32": (i32, [vp, u32, C.c_int32, u32, vp, u64, vp, u64, u32, vp]),
            "uh_fft_c2c_f32": (i32, [vp, u64, u64, vp, u64, u64, u32, u32, vp]),
            "uh_invert_index": (i32, [vp, u32, u32, u32, u32, u32, u32, vp, vp, vp]),
            "uh_prii_rank": (i32, [vp, u32, vp, f32, f32, u32, vp, u32, u32, vp]),
            "uh_sentinel_compress": (i32, [vp, u64, i32, vp, u64, C.POINTER(u64), vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _declared = True
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: status {rc}: {lib().uh_last_error().decode()}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = vp()
        check(lib().uh_malloc(C.byref(p), max(1, self.nbytes)), "uh_malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a) if not a.flags["F_CONTIGUOUS"] else a
        b = cls(a.nbytes)
        check(lib().uh_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes, None), "h2d")
        check(lib().uh_stream_sync(None), "sync")
        return b

    def to_array(self, shape, dtype, order="C") -> np.ndarray:
        out = np.empty(shape, dtype=dtype, order=order)
        check(lib().uh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None), "d2h")
        check(lib().uh_stream_sync(None), "sync")
        return out

    def free(self):
        if self.ptr:
            lib().uh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = vp()
        check(lib().uh_stream_create(C.byref(p)), "stream")
        self.h = p.value

    def sync(self):
        check(lib().uh_stream_sync(self.h), "stream sync")


class Event:
    def __init__(self):
        p = vp()
        check(lib().uh_event_create(C.byref(p)), "event")
        self.h = p.value

    def record(self, stream):
        check(lib().uh_event_record(self.h, stream.h if isinstance(stream, Stream) else stream), "record")

    def elapsed_ms(self, later: "Event") -> float:
        check(lib().uh_event_sync(later.h), "event sync")
        ms = f32(0)
        check(lib().uh_event_elapsed_ms(C.byref(ms), self.h, later.h), "elapsed")
        return float(ms.value)


def gemm(A: DeviceBuffer, B: DeviceBuffer, Cb: DeviceBuffer, M, N, K, lda, ldb, ldc, precision=0,
         alpha=1.0, beta=0.0, stream=None) -> None:
    g = uh_gemm_args(A.ptr, B.ptr, Cb.ptr, lda, ldb, ldc, M, N, K, precision, alpha, beta)
    check(lib().uh_gemm(C.byref(g), stream.h if isinstance(stream, Stream) else stream), "uh_gemm")
