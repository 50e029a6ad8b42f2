"""unicorn_amd — MI355X-native device subtask path of the Unicorn (pm::) task library.

The product is native: the pm:: runtime (libpmlib.so), the gfx950 HIP kernels behind a C-ABI
(libunicorn_hip.so) and the four benchmark apps with C-ABI drivers (libunicorn_apps.so).
This module only loads those libraries with ctypes and offers thin Python handles over the app
drivers for tests and benchmarks. There is no Python or CPU fallback for any GPU path: if a
library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

PM_HIP_MODE_PLAIN = 0
PM_HIP_MODE_RESIDENT = 1
PM_HIP_MODE_BATCHED = 2
DEVICE_CPU, DEVICE_GPU, DEVICE_BOTH = 0, 1, 2

_libs: dict[str, C.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = LIB_DIR / f"lib{name}.so"
    if not path.exists():
        raise NativeLibraryMissing(f"{path} is not built; run `python -m unicorn_amd.build`")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def hip_lib() -> C.CDLL:
    return _load("unicorn_hip")


def pm_lib() -> C.CDLL:
    hip_lib()
    return _load("pmlib")


def apps_lib() -> C.CDLL:
    pm_lib()
    lib = _load("unicorn_apps")
    _declare_apps(lib)
    return lib


class ua_stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "subtasks", "batches", "h2d_bytes", "d2h_bytes", "d2d_bytes", "cache_hits", "evictions",
        "plain_launches", "persistent_launches", "queue_items",
        "ns_queue", "ns_stage", "ns_flush", "ns_wait", "ns_retire", "ns_wake", "ma_copies", "ma_wins",
        "oom_waits", "sentinel_compressed")]


class ua_if_config(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("radius", C.c_int32),
                ("mode", C.c_int32), ("shift", C.c_int32), ("tile", C.c_int32),
                ("weights", C.c_int8 * 1024)]


_declared = False


def _declare_apps(lib: C.CDLL) -> None:
    global _declared
    if _declared:
        return
    vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
    sig = {
        "ua_init": (i32, [i32]), "ua_finalize": (i32, []), "ua_has_gpu": (i32, []),
        "ua_set_hip_mode": (i32, [i32]), "ua_set_hip_compact": (i32, [i32]),
        "ua_set_hip_graphs": (i32, [i32]), "ua_set_hip_depth": (i32, [i32]),
        "ua_set_batch_limit": (i32, [i32]), "ua_set_hbm_budget": (i32, [C.c_uint64]),
        "ua_set_sentinel_threshold": (i32, [C.c_uint64, C.POINTER(C.c_uint64)]),
        "ua_set_split": (i32, [i32, i32, i32, i32]), "ua_set_multi_assign": (i32, [i32]),
        "ua_get_stats": (i32, [C.POINTER(ua_stats)]), "ua_reset_stats": (i32, []),
        "ua_last_error": (C.c_char_p, []),
        "ua_set_kernel_timing": (i32, [i32]),
        "ua_kernel_time": (i32, [C.POINTER(dbl), C.POINTER(C.c_uint64), i32]),
        "ua_device_sync": (i32, []),
        "ua_kernel_time_family": (i32, [i32, C.POINTER(dbl), C.POINTER(C.c_uint64), C.POINTER(dbl), i32]),
        "ua_if_create": (vp, [C.POINTER(ua_if_config)]), "ua_if_set_input": (i32, [vp, vp]),
        "ua_if_run": (i32, [vp, i32, C.POINTER(dbl)]), "ua_if_get_output": (i32, [vp, vp]),
        "ua_if_destroy": (i32, [vp]),
        "ua_mm_create": (vp, [i32, i32, i32]), "ua_mm_set_inputs": (i32, [vp, vp, vp]),
        "ua_mm_run": (i32, [vp, i32, C.POINTER(dbl)]), "ua_mm_get_output": (i32, [vp, vp]),
        "ua_mm_destroy": (i32, [vp]),
        "ua_mt_create": (vp, [C.c_uint64, C.c_uint64, i32]), "ua_mt_set_input": (i32, [vp, vp]),
        "ua_mt_run": (i32, [vp, i32, C.POINTER(dbl)]), "ua_mt_get_output": (i32, [vp, vp]),
        "ua_mt_destroy": (i32, [vp]),
        "ua_ps_create": (vp, [C.c_uint32, C.c_uint32]), "ua_ps_set_input": (i32, [vp, vp]),
        "ua_ps_run": (i32, [vp, i32, C.POINTER(dbl)]), "ua_ps_get_output": (i32, [vp, vp]),
        "ua_ps_destroy": (i32, [vp]),
        "ua_ss_create": (vp, [C.c_uint64, C.c_uint64, vp, vp, vp, vp]), "ua_ss_run": (i32, [vp, i32, C.POINTER(dbl)]),
        "ua_ss_get_output": (i32, [vp, vp]), "ua_ss_destroy": (i32, [vp]),
        "ua_fft_create": (vp, [C.c_uint32, C.c_uint32, C.c_uint32]), "ua_fft_set_input": (i32, [vp, vp]),
        "ua_fft_run": (i32, [vp, i32, C.POINTER(dbl)]), "ua_fft_get_output": (i32, [vp, vp]),
        "ua_fft_destroy": (i32, [vp]),
        "ua_prii_create": (vp, [C.c_char_p, C.c_uint32]), "ua_prii_run": (i32, [vp, i32, C.c_uint32, C.POINTER(dbl)]),
        "ua_prii_get_index": (i32, [vp, vp, C.c_uint64, C.POINTER(C.c_uint64), vp]),
        "ua_prii_get_ranks": (i32, [vp, vp]),
        "ua_prii_info": (i32, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "ua_prii_destroy": (i32, [vp]),
        "ua_webgen": (i32, [C.c_uint32, C.c_uint64, dbl, C.c_uint32, dbl, vp, vp, C.c_uint64,
                            C.POINTER(C.c_uint64)]),
        "ua_webgen_shift": (dbl, [dbl, C.c_uint32, dbl]),
        "ua_pr_create": (vp, [C.c_uint32, vp, vp, C.c_uint32, C.c_float]),
        "ua_pr_run": (i32, [vp, i32, i32, C.POINTER(dbl)]),
        "ua_pr_get_ranks": (i32, [vp, vp]), "ua_pr_edges": (C.c_uint64, [vp]),
        "ua_pr_destroy": (i32, [vp]),
        "ua_pr_block_counts": (i32, [C.c_uint32, vp, vp, vp]),
        "ua_pr_graph_bytes": (C.c_uint64, [C.c_uint32, C.c_uint64, i32]),
        "ua_pr_build_graph": (i32, [C.c_uint32, vp, vp, vp, vp]),
        "ua_pr_set_layout": (None, [i32]), "ua_pr_blocked": (i32, [vp]), "ua_pr_layout": (i32, [vp]),
        "ua_prb_layout": (C.c_uint64, [C.c_uint32, vp, vp, C.c_uint32, C.c_uint32, vp, vp]),
        "ua_webdump_meta": (i32, [C.c_char_p, vp, vp, vp]),
        "ua_webdump_read": (i32, [C.c_char_p, vp, vp, C.c_uint64, vp]),
        "ua_webdump_write": (i32, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp]),
        "ua_pr_create_from_webdump": (vp, [C.c_char_p, C.c_uint32, C.c_float]),
        "ua_prp_create": (vp, [C.c_uint32, vp, vp, C.c_uint32, C.c_float]),
        "ua_prp_create_from_webdump": (vp, [C.c_char_p, C.c_uint32, C.c_float]),
        "ua_prp_run": (i32, [vp, i32, i32, C.POINTER(dbl)]),
        "ua_prp_get_ranks": (i32, [vp, vp]), "ua_prp_destroy": (i32, [vp]),
        "ua_lu_create": (vp, [i32, i32, i32]), "ua_lu_set_matrix": (i32, [vp, vp]),
        "ua_lu_factor": (i32, [vp, i32, C.POINTER(dbl)]), "ua_lu_get_factors": (i32, [vp, vp]),
        "ua_lu_solve": (i32, [vp, vp, i32, i32, i32, vp, vp, C.POINTER(dbl)]),
        "ua_lu_destroy": (i32, [vp]), "ua_lu_prefetch": (i32, [vp]),
        "ua_mm_clone": (vp, [vp]), "ua_if_clone": (vp, [vp]), "ua_pr_clone": (vp, [vp]),
        "ua_mm_fetch": (i32, [vp, vp]), "ua_if_fetch": (i32, [vp, vp]), "ua_pr_fetch": (i32, [vp, vp]),
        "ua_serve_create": (vp, []),
        "ua_serve_add_mm": (i32, [vp, i32, i32, i32, C.c_uint64]),
        "ua_serve_add_if": (i32, [vp, i32, i32, i32, i32, i32, C.c_uint64]),
        "ua_serve_add_pr": (i32, [vp, C.c_uint32, C.c_uint32, i32, C.c_uint64]),
        "ua_serve_run": (i32, [vp, i32, vp, vp, i32, i32, i32, vp, vp, C.POINTER(dbl)]),
        "ua_serve_destroy": (i32, [vp]),
        "ua_serve_dataset": (vp, [vp, i32]), "ua_hash": (C.c_uint64, [vp, C.c_uint64]),
        "ua_mm_get_inputs": (i32, [vp, vp, vp]), "ua_if_get_input": (i32, [vp, vp]),
        "ua_if_get_config": (i32, [vp, C.POINTER(ua_if_config)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _declared = True


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = apps_lib().ua_last_error()
        raise RuntimeError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}")


_runtime_up = False


def init(cpu_stubs: int = 0) -> None:
    """Start the pm:: runtime (idempotent)."""
    global _runtime_up
    if _runtime_up:
        return
    _check(apps_lib().ua_init(cpu_stubs), "pmInitialize")
    _runtime_up = True


def has_gpu() -> bool:
    init()
    return bool(apps_lib().ua_has_gpu())


def set_hip_mode(mode: int, compact: bool = False, graphs: bool = True) -> None:
    init()
    lib = apps_lib()
    _check(lib.ua_set_hip_mode(mode), "pmSetHipStubMode")
    _check(lib.ua_set_hip_compact(1 if compact else 0), "pmSetHipOptimalView")
    _check(lib.ua_set_hip_graphs(1 if graphs else 0), "pmSetHipGraphs")


def set_hbm_budget(nbytes: int) -> None:
    """HBM budget of the runtime in bytes (0 = the default)."""
    _check(apps_lib().ua_set_hbm_budget(nbytes), "ua_set_hbm_budget")


def set_sentinel_threshold(nbytes: int) -> int:
    """Size above which PLAIN-mode reduction outputs go back sentinel-compressed
    (pmSetSentinelCompressionThreshold; the default is 1 GiB). Returns the value replaced."""
    init()
    prev = C.c_uint64(0)
    _check(apps_lib().ua_set_sentinel_threshold(nbytes, C.byref(prev)), "pmSetSentinelCompressionThreshold")
    return int(prev.value)


def set_split(cpu: bool = True, gpu: bool = False, cpu_pieces: int = 0, gpu_pieces: int = 0) -> None:
    """Subtask splitting of imageFiltering / matrixMultiply / fft tasks submitted afterwards: the
    canSplitCpuSubtasks / canSplitGpuSubtasks flags and the pieces per split subtask (0 = one per
    stub of the splitting device type). The defaults are the reference apps' settings."""
    init()
    _check(apps_lib().ua_set_split(int(cpu), int(gpu), cpu_pieces, gpu_pieces), "ua_set_split")


def set_multi_assign(on: bool = True) -> None:
    """multiAssignEnabled of imageFiltering / matrixMultiply tasks submitted afterwards."""
    init()
    _check(apps_lib().ua_set_multi_assign(int(on)), "ua_set_multi_assign")


def stats() -> dict:
    s = ua_stats()
    _check(apps_lib().ua_get_stats(C.byref(s)), "stats")
    return {n: int(getattr(s, n)) for n, _ in ua_stats._fields_}


def reset_stats() -> None:
    _check(apps_lib().ua_reset_stats(), "reset stats")


def set_kernel_timing(on: bool, every: int = 1) -> None:
    """Events around the persistent kernels of every `every`-th BATCHED flush (all when 1)."""
    _check(apps_lib().ua_set_kernel_timing(every if on else 0), "pmSetHipKernelTimingEvery")


def kernel_time(reset: bool = False):
    """(milliseconds, flushes) of BATCHED flushes since the last reset, timed on their stream."""
    ms, n = C.c_double(0), C.c_uint64(0)
    _check(apps_lib().ua_kernel_time(C.byref(ms), C.byref(n), 1 if reset else 0), "pmGetHipKernelTime")
    return ms.value, int(n.value)


# kernel families of the batched flushes (UH_FAMILY_* of unicorn_hip.h)
FAMILY_FILTER, FAMILY_GEMM, FAMILY_PAGERANK, FAMILY_TRSM, FAMILY_PAGERANK_BLOCKED, FAMILY_PAGERANK_BINNED = 0, 1, 2, 3, 4, 5


def kernel_time_family(family: int, reset: bool = False):
    """(milliseconds, flushes, algorithmic work) of the timed BATCHED flushes that launched only
    `family`'s kernel: work is flop issued (GEMM), bytes moved (filter) or 0."""
    ms, n, w = C.c_double(0), C.c_uint64(0), C.c_double(0)
    _check(apps_lib().ua_kernel_time_family(family, C.byref(ms), C.byref(n), C.byref(w), 1 if reset else 0),
           "pmGetHipKernelTimeByFamily")
    return ms.value, int(n.value), w.value


def device_sync() -> None:
    _check(apps_lib().ua_device_sync(), "device sync")


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class ImageFilter:
    """imageFiltering task over an H x W RGB image (rows stored bottom-up, as the reference)."""

    def __init__(self, width: int, height: int, weights: np.ndarray, radius: int, mode: int = 0,
                 shift: int = 4, tile: int = 2048):
        init()
        cfg = ua_if_config()
        cfg.width, cfg.height, cfg.radius, cfg.mode, cfg.shift, cfg.tile = width, height, radius, mode, shift, tile
        w = np.asarray(weights, dtype=np.int8).reshape(-1)
        if w.size != (2 * radius + 1) ** 2:
            raise ValueError("weights must hold (2r+1)^2 values")
        for i, v in enumerate(w):
            cfg.weights[i] = int(v)
        self.shape = (height, width, 3)
        self._h = apps_lib().ua_if_create(C.byref(cfg))
        if not self._h:
            _check(-1, "ua_if_create")

    def set_input(self, image: np.ndarray) -> None:
        image = np.ascontiguousarray(image, dtype=np.uint8)
        if image.shape != self.shape:
            raise ValueError(f"image shape {image.shape} != {self.shape}")
        _check(apps_lib().ua_if_set_input(self._h, _ptr(image)), "ua_if_set_input")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_if_run(self._h, device, C.byref(t)), "imageFilter task")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint8)
        _check(apps_lib().ua_if_get_output(self._h, _ptr(out)), "ua_if_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_if_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def webgen(pages: int, seed: int, mean_out: float = 10.0, max_out: int = 100, dangling: float = 0.05):
    """Synthetic power-law web graph as an out-link CSR (row_ptr, col) with 0-based targets."""
    lib = apps_lib()
    row_ptr = np.empty(pages + 1, dtype=np.uint32)
    edges = C.c_uint64(0)
    _check(lib.ua_webgen(pages, seed, mean_out, max_out, dangling, _ptr(row_ptr), None, 0, C.byref(edges)),
           "ua_webgen")
    col = np.empty(max(1, edges.value), dtype=np.uint32)
    _check(lib.ua_webgen(pages, seed, mean_out, max_out, dangling, _ptr(row_ptr), _ptr(col), col.size,
                         C.byref(edges)), "ua_webgen")
    return row_ptr, col[:edges.value]


class PageRank:
    """pageRank over an out-link CSR; subtasks own `pages_per_subtask` destination pages."""

    def __init__(self, row_ptr: np.ndarray, col: np.ndarray, pages_per_subtask: int, damping: float = 0.85):
        init()
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col = np.ascontiguousarray(col if col.size else np.zeros(1, np.uint32), dtype=np.uint32)
        self.pages = self.row_ptr.size - 1
        self._h = apps_lib().ua_pr_create(self.pages, _ptr(self.row_ptr), _ptr(self.col), pages_per_subtask, damping)
        if not self._h:
            _check(-1, "ua_pr_create")

    @classmethod
    def from_webdump(cls, path: str, pages_per_subtask: int, damping: float = 0.85) -> "PageRank":
        """pageRank over a web dump in the reference's on-disk format (see webdump_write)."""
        init()
        self = cls.__new__(cls)
        pages, max_out, per_file = webdump_meta(path)
        self.pages = pages
        self.row_ptr = self.col = None
        self._h = apps_lib().ua_pr_create_from_webdump(str(path).encode(), pages_per_subtask, damping)
        if not self._h:
            _check(-1, "ua_pr_create_from_webdump")
        return self

    @property
    def blocked(self) -> bool:
        """True if the graph is held in the source-blocked layout."""
        return bool(apps_lib().ua_pr_blocked(self._h))

    @property
    def layout(self) -> str:
        """'csr', 'blocked' or 'binned': the graph layout this instance holds."""
        return PR_LAYOUTS[apps_lib().ua_pr_layout(self._h)]

    def run(self, iterations: int, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_pr_run(self._h, iterations, device, C.byref(t)), "pageRank tasks")
        return t.value

    def ranks(self) -> np.ndarray:
        out = np.empty(self.pages, dtype=np.float32)
        _check(apps_lib().ua_pr_get_ranks(self._h, _ptr(out)), "ua_pr_get_ranks")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_pr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PageRankPush:
    """pageRank in the reference's decomposition: source-page subtasks push into private
    full-length vectors that the built-in REDUCE_ADD folds (on the device in the HIP stub's
    resident modes). Build from an out-link CSR or, with from_webdump, from a web dump."""

    def __init__(self, row_ptr: np.ndarray, col: np.ndarray, pages_per_subtask: int, damping: float = 0.85):
        init()
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col = np.ascontiguousarray(col if col.size else np.zeros(1, np.uint32), dtype=np.uint32)
        self.pages = row_ptr.size - 1
        self._h = apps_lib().ua_prp_create(self.pages, _ptr(row_ptr), _ptr(col), pages_per_subtask, damping)
        if not self._h:
            _check(-1, "ua_prp_create")

    @classmethod
    def from_webdump(cls, path: str, pages_per_subtask: int, damping: float = 0.85) -> "PageRankPush":
        init()
        self = cls.__new__(cls)
        self.pages = webdump_meta(path)[0]
        self._h = apps_lib().ua_prp_create_from_webdump(str(path).encode(), pages_per_subtask, damping)
        if not self._h:
            _check(-1, "ua_prp_create_from_webdump")
        return self

    def run(self, iterations: int, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_prp_run(self._h, iterations, device, C.byref(t)), "pageRank push tasks")
        return t.value

    def ranks(self) -> np.ndarray:
        out = np.empty(self.pages, dtype=np.float32)
        _check(apps_lib().ua_prp_get_ranks(self._h, _ptr(out)), "ua_prp_get_ranks")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            apps_lib().ua_prp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def webdump_meta(path: str):
    """(pages, max out-links per page, pages per file) of a web dump."""
    v = [C.c_uint32(0) for _ in range(3)]
    _check(apps_lib().ua_webdump_meta(str(path).encode(), *[C.byref(x) for x in v]), "ua_webdump_meta")
    return tuple(x.value for x in v)


def webdump_read(path: str):
    """The out-link CSR (row_ptr, col; 0-based targets) of a web dump in the reference's format."""
    pages, _, _ = webdump_meta(path)
    lib = apps_lib()
    row_ptr = np.empty(pages + 1, dtype=np.uint32)
    edges = C.c_uint64(0)
    _check(lib.ua_webdump_read(str(path).encode(), _ptr(row_ptr), None, 0, C.byref(edges)), "ua_webdump_read")
    col = np.empty(max(1, edges.value), dtype=np.uint32)
    _check(lib.ua_webdump_read(str(path).encode(), _ptr(row_ptr), _ptr(col), col.size, C.byref(edges)),
           "ua_webdump_read")
    return row_ptr, col[:edges.value]


def webdump_write(path: str, row_ptr: np.ndarray, col: np.ndarray, max_out: int, pages_per_file: int) -> None:
    """Writes an out-link CSR as a web dump in the reference's format: <path>/web_dump_metadata
    and <path>/web/page_<n> files of fixed records (count, 1-based targets, zero padding)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col if col.size else np.zeros(1, np.uint32), dtype=np.uint32)
    _check(apps_lib().ua_webdump_write(str(path).encode(), row_ptr.size - 1, max_out, pages_per_file,
                                       _ptr(row_ptr), _ptr(col)), "ua_webdump_write")


PR_LAYOUTS = ("csr", "blocked", "binned")


def pr_set_layout(layout) -> None:
    """Layout of the pageRank graphs created from now on: 'binned' (the default), 'blocked' or
    'csr' (True / False mean 'blocked' / 'csr'); a layout that cannot hold a graph falls back to
    the next one (binned -> blocked -> csr)."""
    if layout is True:
        layout = "blocked"
    elif layout is False:
        layout = "csr"
    apps_lib().ua_pr_set_layout(PR_LAYOUTS.index(layout))


def pr_binned_layout(row_ptr: np.ndarray, col: np.ndarray, pages_per_subtask: int, cus: int = 256):
    """The binned graph address space built from an out-link CSR, as numpy views: a dict with
    units, runs, slots, subtasks and the arrays deg, unit_start, bin_start, run_ptr, run_off,
    run_pos, dst16, src16; None when the layout cannot hold the graph."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col if col.size else np.zeros(1, np.uint32), dtype=np.uint32)
    n = row_ptr.size - 1
    L = apps_lib()
    info = np.zeros(12, dtype=np.uint64)
    nbytes = int(L.ua_prb_layout(n, _ptr(row_ptr), _ptr(col), pages_per_subtask, cus, None, _ptr(info)))
    if nbytes == 0:
        return None
    g = np.zeros(nbytes, dtype=np.uint8)
    if int(L.ua_prb_layout(n, _ptr(row_ptr), _ptr(col), pages_per_subtask, cus, _ptr(g), _ptr(info))) != nbytes:
        raise RuntimeError("ua_prb_layout: build failed")
    K, R, S, subtasks = (int(v) for v in info[:4])
    off = [int(v) for v in info[4:]]
    u32 = lambda o, n_: g[o:o + 4 * n_].view(np.uint32)
    u16 = lambda o, n_: g[o:o + 2 * n_].view(np.uint16)
    return {"units": K, "runs": R, "slots": S, "subtasks": subtasks,
            "deg": u32(off[0], n), "unit_start": u32(off[1], K + 1), "bin_start": u32(off[2], K + 1),
            "run_ptr": u32(off[3], K + 1), "run_off": u32(off[4], R + 1), "run_pos": u32(off[5], R),
            "dst16": u16(off[6], S), "src16": u16(off[7], int(row_ptr[-1]))}


# Row group size and source block count of the source-blocked pageRank layout
# (UH_PR_GROUP_ROWS and UH_PR_BLOCKS of unicorn_hip.h).
PR_GROUP_ROWS = 2048
PR_BLOCKS = 8


def pr_graph_layout(row_ptr: np.ndarray, col: np.ndarray, blocked: bool = True):
    """The graph address space that PageRank builds from an out-link CSR, as numpy views:
    (in_row_ptr, outdeg, in_col, bseg, bcnt); bseg and bcnt are None for the plain CSR (or when
    a page has more than 255 in-links from one source block)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col if col.size else np.zeros(1, np.uint32), dtype=np.uint32)
    n = row_ptr.size - 1
    e = int(row_ptr[-1])
    GR, NB = PR_GROUP_ROWS, PR_BLOCKS
    groups = (n + GR - 1) // GR
    L = apps_lib()
    bcnt = np.zeros(groups * NB * GR, dtype=np.uint8)
    if blocked and not L.ua_pr_block_counts(n, _ptr(row_ptr), _ptr(col), _ptr(bcnt)):
        blocked = False
    nbytes = int(L.ua_pr_graph_bytes(n, e, 1 if blocked else 0))
    g = np.zeros((nbytes + 3) // 4, dtype=np.uint32)
    _check(L.ua_pr_build_graph(n, _ptr(row_ptr), _ptr(col), _ptr(bcnt) if blocked else None, _ptr(g)),
           "ua_pr_build_graph")
    rp, deg, icol = g[:n + 1], g[n + 1:2 * n + 1], g[2 * n + 1:2 * n + 1 + e]
    if not blocked:
        return rp, deg, icol, None, None
    bseg = g[2 * n + 1 + e:2 * n + 1 + e + groups * NB + 1]
    off = (2 * n + 1 + e + groups * NB + 1) * 4
    return rp, deg, icol, bseg, g.view(np.uint8)[off:off + groups * NB * GR]


class LU:
    """luDecomposition (blocked, 3 tasks per block step) and the refinement solve."""

    def __init__(self, n: int, block: int = 512, precision: int = 0):
        init()
        self.n = n
        self._h = apps_lib().ua_lu_create(n, block, precision)
        if not self._h:
            _check(-1, "ua_lu_create")

    def set_matrix(self, A_colmajor: np.ndarray) -> None:
        a = np.asfortranarray(A_colmajor, dtype=np.float32)
        assert a.shape == (self.n, self.n)
        _check(apps_lib().ua_lu_set_matrix(self._h, a.ctypes.data), "ua_lu_set_matrix")

    def prefetch(self) -> None:
        _check(apps_lib().ua_lu_prefetch(self._h), "ua_lu_prefetch")

    def factor(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_lu_factor(self._h, device, C.byref(t)), "luDecomposition tasks")
        return t.value

    def factors(self) -> np.ndarray:
        out = np.empty((self.n, self.n), dtype=np.float32, order="F")
        _check(apps_lib().ua_lu_get_factors(self._h, out.ctypes.data), "ua_lu_get_factors")
        return out

    def solve(self, b: np.ndarray, max_steps: int = 5, device: int = DEVICE_GPU):
        b = np.asfortranarray(b, dtype=np.float64)
        nrhs = b.shape[1]
        x = np.empty_like(b, order="F")
        info = (C.c_double * 4)()
        t = C.c_double(0.0)
        _check(apps_lib().ua_lu_solve(self._h, b.ctypes.data, nrhs, max_steps, device, x.ctypes.data, info,
                                      C.byref(t)), "refinement solve")
        return x, {"steps": int(info[0]), "backward_error": info[1], "backward_error_first": info[2],
                   "a_inf": info[3], "seconds": t.value}

    def close(self) -> None:
        if self._h:
            apps_lib().ua_lu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MatMul:
    """matrixMultiply task: C = A B, n x n float32, (n/block)^2 block subtasks."""

    def __init__(self, n: int, block: int = 512, precision: int = 0):
        init()
        self.n = n
        self._h = apps_lib().ua_mm_create(n, block, precision)
        if not self._h:
            _check(-1, "ua_mm_create")

    def set_inputs(self, A: np.ndarray, B: np.ndarray) -> None:
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        assert A.shape == (self.n, self.n) and B.shape == (self.n, self.n)
        _check(apps_lib().ua_mm_set_inputs(self._h, _ptr(A), _ptr(B)), "ua_mm_set_inputs")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_mm_run(self._h, device, C.byref(t)), "matrixMultiply task")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty((self.n, self.n), dtype=np.float32)
        _check(apps_lib().ua_mm_get_output(self._h, _ptr(out)), "ua_mm_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_mm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MatTranspose:
    """matrixTranspose task: int32 rows x cols -> cols x rows, one subtask per square block of
    min(rows, cols, 2048); in place uses one READ_WRITE address space."""

    def __init__(self, rows: int, cols: int, inplace: bool = False):
        init()
        self.rows, self.cols = rows, cols
        self._h = apps_lib().ua_mt_create(rows, cols, int(inplace))
        if not self._h:
            _check(-1, "ua_mt_create")

    def set_input(self, M: np.ndarray) -> None:
        M = np.ascontiguousarray(M, dtype=np.int32)
        assert M.shape == (self.rows, self.cols)
        _check(apps_lib().ua_mt_set_input(self._h, _ptr(M)), "ua_mt_set_input")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_mt_run(self._h, device, C.byref(t)), "matrixTranspose task")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty((self.cols, self.rows), dtype=np.int32)
        _check(apps_lib().ua_mt_get_output(self._h, _ptr(out)), "ua_mt_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_mt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PrefixSum:
    """prefixSum task pair: inclusive int32 prefix sum (wrap-around) of n elements in chunks."""

    def __init__(self, n: int, elems_per_subtask: int = 1_000_000):
        init()
        self.n = n
        self._h = apps_lib().ua_ps_create(n, elems_per_subtask)
        if not self._h:
            _check(-1, "ua_ps_create")

    def set_input(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.int32)
        assert x.shape == (self.n,)
        _check(apps_lib().ua_ps_set_input(self._h, _ptr(x)), "ua_ps_set_input")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_ps_run(self._h, device, C.byref(t)), "prefixSum tasks")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(apps_lib().ua_ps_get_output(self._h, _ptr(out)), "ua_ps_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_ps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SparseMatMul:
    """sparseSolver task: C = A B with A sparse (COO sorted by row, then column), B dense, n x n."""

    def __init__(self, n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray, B: np.ndarray):
        init()
        self.n = n
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        self._h = apps_lib().ua_ss_create(n, rows.size, _ptr(rows), _ptr(cols), _ptr(vals), _ptr(B))
        if not self._h:
            _check(-1, "ua_ss_create")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_ss_run(self._h, device, C.byref(t)), "sparseSolver task")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty((self.n, self.n), dtype=np.float32)
        _check(apps_lib().ua_ss_get_output(self._h, _ptr(out)), "ua_ss_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_ss_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Fft2D:
    """fft tasks: the forward, unnormalised 2-D DFT of a rows x cols complex64 matrix (a row task,
    then a column task of rows_per_subtask rows / columns per subtask; 0 = 512)."""

    def __init__(self, rows: int, cols: int, rows_per_subtask: int = 0):
        init()
        self.shape = (rows, cols)
        self._h = apps_lib().ua_fft_create(rows, cols, rows_per_subtask)
        if not self._h:
            _check(-1, "ua_fft_create")

    def set_input(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        if x.shape != self.shape:
            raise ValueError(f"input shape {x.shape} != {self.shape}")
        _check(apps_lib().ua_fft_set_input(self._h, _ptr(x)), "ua_fft_set_input")

    def run(self, device: int = DEVICE_GPU) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_fft_run(self._h, device, C.byref(t)), "fft tasks")
        return t.value

    def output(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.complex64)
        _check(apps_lib().ua_fft_get_output(self._h, _ptr(out)), "ua_fft_get_output")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_fft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PageRankInvertedIndex:
    """pageRankInvertedIndex tasks over a web dump directory: the index inversion (with data
    redistribution), then pageRank iterations over the inverted index."""

    def __init__(self, directory: str, pages_per_subtask: int = 0):
        init()
        self._h = apps_lib().ua_prii_create(str(directory).encode(), pages_per_subtask)
        if not self._h:
            _check(-1, "ua_prii_create")
        p, per, n = C.c_uint32(), C.c_uint32(), C.c_uint32()
        apps_lib().ua_prii_info(self._h, C.byref(p), C.byref(per), C.byref(n))
        self.pages, self.pages_per_subtask, self.subtasks = p.value, per.value, n.value

    def run(self, device: int = DEVICE_GPU, iterations: int = 0) -> float:
        t = C.c_double(0.0)
        _check(apps_lib().ua_prii_run(self._h, device, iterations, C.byref(t)), "pageRankInvertedIndex tasks")
        return t.value

    def index(self):
        """(entries [n x 3] uint32: src, dest, src out-links; inlink counts per subtask)."""
        n = C.c_uint64()
        counts = np.zeros(self.subtasks, np.uint32)
        _check(apps_lib().ua_prii_get_index(self._h, None, 0, C.byref(n), _ptr(counts)), "ua_prii_get_index")
        e = np.zeros((n.value, 3), np.uint32)
        _check(apps_lib().ua_prii_get_index(self._h, _ptr(e), n.value, C.byref(n), None), "ua_prii_get_index")
        return e, counts

    def ranks(self) -> np.ndarray:
        out = np.empty(self.pages, np.float32)
        _check(apps_lib().ua_prii_get_ranks(self._h, _ptr(out)), "ua_prii_get_ranks")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_prii_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving driver over a catalogue of resident datasets (ua_serve_* of libunicorn_apps.so)."""

    def __init__(self):
        self._h = apps_lib().ua_serve_create()
        self.kinds: list[tuple] = []

    def _add(self, ds: int, what: str, desc: tuple) -> int:
        if ds < 0:
            _check(-1, what)
        self.kinds.append(desc)
        return ds

    def add_mm(self, n: int, block: int = 512, precision: int = 0, seed: int = 1) -> int:
        return self._add(apps_lib().ua_serve_add_mm(self._h, n, min(block, n), precision, seed),
                         "ua_serve_add_mm", ("mm", n))

    def add_if(self, side: int, radius: int = 2, mode: int = 1, shift: int = 4, tile: int = 2048,
               seed: int = 1) -> int:
        return self._add(apps_lib().ua_serve_add_if(self._h, side, radius, mode, shift, tile, seed),
                         "ua_serve_add_if", ("if", side))

    def add_pr(self, pages: int, pages_per_subtask: int = 1_000_000, iterations: int = 20,
               seed: int = 1) -> int:
        return self._add(apps_lib().ua_serve_add_pr(self._h, pages, pages_per_subtask, iterations, seed),
                         "ua_serve_add_pr", ("pr", pages))

    def run(self, dataset: np.ndarray, arrival_s: np.ndarray, slots: int, device: int = DEVICE_GPU,
            fetch: bool = True, checksums: bool = False):
        """Serve the request stream; returns (latency_ms, checksums or None, wall seconds)."""
        ds = np.ascontiguousarray(dataset, dtype=np.int32)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        n = len(ds)
        lat = np.zeros(n, dtype=np.float64)
        cs = np.zeros(n, dtype=np.uint64) if checksums else None
        wall = C.c_double(0.0)
        _check(apps_lib().ua_serve_run(self._h, n, _ptr(ds), _ptr(arr), slots, device, 1 if fetch else 0,
                                       _ptr(lat), _ptr(cs) if cs is not None else None, C.byref(wall)),
               "serving run")
        return lat, cs, wall.value

    def _base(self, ds: int):
        h = apps_lib().ua_serve_dataset(self._h, ds)
        if not h:
            _check(-1, "ua_serve_dataset")
        return h

    def mm_inputs(self, ds: int):
        """(A, B) of a matrixMultiply dataset."""
        n = self.kinds[ds][1]
        A = np.empty((n, n), dtype=np.float32)
        B = np.empty((n, n), dtype=np.float32)
        _check(apps_lib().ua_mm_get_inputs(self._base(ds), _ptr(A), _ptr(B)), "ua_mm_get_inputs")
        return A, B

    def if_input(self, ds: int):
        """(image rows bottom-up, weights, radius, mode, shift) of an imageFiltering dataset."""
        cfg = ua_if_config()
        h = self._base(ds)
        _check(apps_lib().ua_if_get_config(h, C.byref(cfg)), "ua_if_get_config")
        img = np.empty((cfg.height, cfg.width, 3), dtype=np.uint8)
        _check(apps_lib().ua_if_get_input(h, _ptr(img)), "ua_if_get_input")
        d = 2 * cfg.radius + 1
        w = np.array([cfg.weights[i] for i in range(d * d)], dtype=np.int8)
        return img, w, cfg.radius, cfg.mode, cfg.shift

    def run_alone(self, ds: int, device: int = DEVICE_GPU, iterations: int = 20) -> np.ndarray:
        """Runs one task of dataset `ds` on the instance that owns its inputs (which no request
        uses) and returns its output: C, the filtered image, or the ranks."""
        kind, size = self.kinds[ds]
        h = self._base(ds)
        L = apps_lib()
        t = C.c_double(0.0)
        if kind == "mm":
            _check(L.ua_mm_run(h, device, C.byref(t)), "matrixMultiply task")
            out = np.empty((size, size), dtype=np.float32)
            _check(L.ua_mm_get_output(h, _ptr(out)), "ua_mm_get_output")
        elif kind == "if":
            _check(L.ua_if_run(h, device, C.byref(t)), "imageFilter task")
            out = np.empty((size, size, 3), dtype=np.uint8)
            _check(L.ua_if_get_output(h, _ptr(out)), "ua_if_get_output")
        else:
            _check(L.ua_pr_run(h, iterations, device, C.byref(t)), "pageRank tasks")
            out = np.empty(size, dtype=np.float32)
            _check(L.ua_pr_get_ranks(h, _ptr(out)), "ua_pr_get_ranks")
        return out

    def close(self) -> None:
        if self._h:
            apps_lib().ua_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def output_hash(a: np.ndarray) -> int:
    """The hash that the serving driver reports for a task output (ua_hash)."""
    a = np.ascontiguousarray(a)
    return int(apps_lib().ua_hash(a.ctypes.data, a.nbytes))


def serving_stream(seed: int, n: int, rate_per_s: float, catalogue: dict, mix=(1, 1, 1)):
    """Seeded request stream of the serving mix: Poisson arrivals at `rate_per_s` (0 = all at
    t = 0) and a class per request drawn with weights `mix` over (matrixMultiply, imageFiltering,
    pageRank). Sizes follow the mix's distributions: matrixMultiply N log-uniform, imageFiltering
    side uniform, pageRank pages uniform; each is snapped to the nearest size of `catalogue`
    ({"mm": [(N, ds)], "if": [(side, ds)], "pr": [(pages, ds)]}, nearest in log scale for N).
    Returns (dataset ids, arrival seconds)."""
    rng = np.random.default_rng(seed)
    classes = [k for k in ("mm", "if", "pr")]
    w = np.asarray(mix, dtype=np.float64)
    w = w / w.sum()
    ds = np.empty(n, dtype=np.int32)
    for i in range(n):
        k = classes[rng.choice(3, p=w)]
        entries = catalogue[k]
        sizes = np.array([e[0] for e in entries], dtype=np.float64)
        lo, hi = sizes.min(), sizes.max()
        if k == "mm":
            v = np.exp(rng.uniform(np.log(lo), np.log(hi)))
            j = int(np.argmin(np.abs(np.log(sizes) - np.log(v))))
        else:
            v = rng.uniform(lo, hi)
            j = int(np.argmin(np.abs(sizes - v)))
        ds[i] = entries[j][1]
    if rate_per_s > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, n))
    else:
        arrivals = np.zeros(n)
    return ds, arrivals
