// Device management, memory, streams/events/graphs, staging kernels and the batch work queue
// of libunicorn_hip.so. Replaces the reference's CUDA runtime wrapper
// (source/code/src/pmCudaInterface.cu:69-378) for one MI355X.
#include "uh_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>

namespace uh {

static thread_local char t_err[512] = "";
std::atomic<uint64_t> g_plain_launches{0};
std::atomic<uint64_t> g_persistent_launches{0};
std::atomic<uint64_t> g_items{0};

int fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_err, sizeof(t_err), fmt, ap);
    va_end(ap);
    return status;
}

int check_hip(hipError_t err, const char* what) {
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
        return fail(UH_ERR_OOM, "%s: %s", what, hipGetErrorString(err));
    return fail(UH_ERR_HIP, "%s: %s", what, hipGetErrorString(err));
}

int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
        cus = p.multiProcessorCount;
    }
    return cus;
}

Batch* as_batch(void* stream) {
    uintptr_t v = (uintptr_t)stream;
    if (v & 1u) return (Batch*)(v & ~(uintptr_t)1);
    return nullptr;
}

hipStream_t plain_stream_of(void* stream, int* rc) {
    *rc = UH_OK;
    if (Batch* b = as_batch(stream)) {
        *rc = uh_batch_flush(stream, 0);
        return b->stream;
    }
    return (hipStream_t)stream;
}

void record_item(Batch* b, Family f, const void* desc, size_t bytes, uint32_t tiles,
                 uint32_t kind_bits) {
    const uint8_t* p = (const uint8_t*)desc;
    b->kinds[f] |= kind_bits;
    b->items[f].insert(b->items[f].end(), p, p + bytes);
    b->tiles[f].push_back(tiles);
}

// ---- staging kernels ----

// One workgroup per record; 16-byte vector copies when both ends are aligned.
__global__ void __launch_bounds__(256) k_copy_records(uint8_t* dst, const uint8_t* src,
                                                       const uint64_t* src_off,
                                                       const uint64_t* dst_off,
                                                       const uint64_t* len) {
    const uint32_t r = blockIdx.x;
    const uint8_t* s = src + src_off[r];
    uint8_t* d = dst + dst_off[r];
    const uint64_t n = len[r];
    const uint64_t chunk = (uint64_t)blockIdx.y * 65536u;
    if (chunk >= n) return;
    const uint64_t end = (chunk + 65536u < n) ? chunk + 65536u : n;
    if ((((uintptr_t)s | (uintptr_t)d | chunk) & 15u) == 0) {
        const uint64_t nv = (end - chunk) >> 4;
        const uint4* sv = (const uint4*)(s + chunk);
        uint4* dv = (uint4*)(d + chunk);
        for (uint64_t i = threadIdx.x; i < nv; i += blockDim.x) dv[i] = sv[i];
        for (uint64_t i = chunk + (nv << 4) + threadIdx.x; i < end; i += blockDim.x) d[i] = s[i];
    } else {
        for (uint64_t i = chunk + threadIdx.x; i < end; i += blockDim.x) d[i] = s[i];
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* dst, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

}  // namespace uh

using namespace uh;

extern "C" {

const char* uh_last_error(void) { return t_err; }

int uh_device_count(int* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; return fail(UH_ERR_NODEV, "hipGetDeviceCount: %s", hipGetErrorString(e)); }
    *count = n;
    return n > 0 ? UH_OK : fail(UH_ERR_NODEV, "no HIP device");
}

int uh_set_device(int device) { UH_HIP(hipSetDevice(device)); return UH_OK; }

int uh_device_props(int device, char* name, size_t name_len, size_t* total_mem, int* cu,
                    int* clock_khz, int* l2_bytes, int* lds_per_block) {
    hipDeviceProp_t p;
    UH_HIP(hipGetDeviceProperties(&p, device));
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (cu) *cu = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (l2_bytes) *l2_bytes = p.l2CacheSize;
    if (lds_per_block) *lds_per_block = (int)p.sharedMemPerBlock;
    return UH_OK;
}

int uh_mem_info(size_t* free_bytes, size_t* total_bytes) {
    UH_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return UH_OK;
}

int uh_device_sync(void) { UH_HIP(hipDeviceSynchronize()); return UH_OK; }

int uh_malloc(void** ptr, size_t bytes) {
    *ptr = nullptr;
    if (!bytes) return UH_OK;
    UH_HIP(hipMalloc(ptr, bytes));
    return UH_OK;
}

int uh_free(void* ptr) { if (ptr) UH_HIP(hipFree(ptr)); return UH_OK; }

int uh_host_alloc(void** ptr, size_t bytes) {
    *ptr = nullptr;
    if (!bytes) return UH_OK;
    UH_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return UH_OK;
}

int uh_host_free(void* ptr) { if (ptr) UH_HIP(hipHostFree(ptr)); return UH_OK; }

static int copy_kind(const char* what, void* dst, const void* src, size_t bytes, void* stream,
                     hipMemcpyKind kind) {
    if (!bytes) return UH_OK;
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    if (e != hipSuccess) return check_hip(e, what);
    return UH_OK;
}

int uh_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_kind("hipMemcpyAsync(H2D)", dst, src, bytes, stream, hipMemcpyHostToDevice);
}
int uh_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_kind("hipMemcpyAsync(D2H)", dst, src, bytes, stream, hipMemcpyDeviceToHost);
}
int uh_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_kind("hipMemcpyAsync(D2D)", dst, src, bytes, stream, hipMemcpyDeviceToDevice);
}

int uh_memcpy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                 size_t rows, int kind, void* stream) {
    if (!width || !rows) return UH_OK;
    if (width > dpitch || width > spitch) return fail(UH_ERR_ARG, "uh_memcpy_2d: width > pitch");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    UH_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, k, s));
    return UH_OK;
}

int uh_fill_u32(void* dst, uint32_t value, size_t n, void* stream) {
    if (!n) return UH_OK;
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    UH_LAUNCH(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s,
                       (uint32_t*)dst, value, n);
    UH_LAUNCH_CHECK("k_fill_u32");
    g_plain_launches++;
    return UH_OK;
}

int uh_stream_create(void** stream) {
    hipStream_t s;
    UH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return UH_OK;
}
int uh_stream_destroy(void* stream) { if (stream) UH_HIP(hipStreamDestroy((hipStream_t)stream)); return UH_OK; }
int uh_stream_sync(void* stream) {
    Batch* b = as_batch(stream);
    UH_HIP(hipStreamSynchronize(b ? b->stream : (hipStream_t)stream));
    return UH_OK;
}
int uh_stream_query(void* stream) {
    Batch* b = as_batch(stream);
    hipError_t e = hipStreamQuery(b ? b->stream : (hipStream_t)stream);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return check_hip(e, "hipStreamQuery") ? -1 : -1;
}
int uh_stream_wait_event(void* stream, void* event) {
    Batch* b = as_batch(stream);
    UH_HIP(hipStreamWaitEvent(b ? b->stream : (hipStream_t)stream, (hipEvent_t)event, 0));
    return UH_OK;
}
int uh_event_create(void** event) {
    hipEvent_t e;
    UH_HIP(hipEventCreate(&e));
    *event = e;
    return UH_OK;
}
int uh_event_destroy(void* event) { if (event) UH_HIP(hipEventDestroy((hipEvent_t)event)); return UH_OK; }
int uh_event_record(void* event, void* stream) {
    Batch* b = as_batch(stream);
    UH_HIP(hipEventRecord((hipEvent_t)event, b ? b->stream : (hipStream_t)stream));
    return UH_OK;
}
int uh_event_sync(void* event) { UH_HIP(hipEventSynchronize((hipEvent_t)event)); return UH_OK; }
int uh_event_query(void* event) {
    hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    check_hip(e, "hipEventQuery");
    return -1;
}
int uh_event_elapsed_ms(float* ms, void* start, void* stop) {
    UH_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return UH_OK;
}

int uh_graph_begin_capture(void* stream) {
    UH_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return UH_OK;
}
int uh_graph_end_capture(void* stream, void** graph_exec) {
    hipGraph_t g;
    UH_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ge;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
    *graph_exec = ge;
    return UH_OK;
}
int uh_graph_launch(void* graph_exec, void* stream) {
    UH_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return UH_OK;
}
int uh_graph_destroy(void* graph_exec) {
    if (graph_exec) UH_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return UH_OK;
}

int uh_launch_kernel(const void* func, int gx, int gy, int gz, int bx, int by, int bz,
                     size_t shared_mem, void* stream, void** args) {
    if (gx <= 0 || gy <= 0 || gz <= 0 || bx <= 0 || by <= 0 || bz <= 0 || bx * by * bz > 1024)
        return fail(UH_ERR_ARG, "uh_launch_kernel: bad geometry");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    UH_HIP(hipLaunchKernel(func, dim3(gx, gy, gz), dim3(bx, by, bz), args, shared_mem, s));
    g_plain_launches++;
    return UH_OK;
}

int uh_copy_records(void* dst, const void* src, const uint64_t* src_off, const uint64_t* dst_off,
                    const uint64_t* len, uint32_t n_records, uint64_t max_len, void* stream) {
    if (!n_records || !max_len) return UH_OK;
    uint64_t chunks = (max_len + 65535) / 65536;
    if (chunks > 65535 || n_records > 0x7fffffffu) return fail(UH_ERR_ARG, "uh_copy_records: too large");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_copy_records, dim3(n_records, (unsigned)chunks), dim3(256), 0,
                       s, (uint8_t*)dst, (const uint8_t*)src, src_off, dst_off, len);
    UH_LAUNCH_CHECK("k_copy_records");
    g_plain_launches++;
    return UH_OK;
}

// ---- batching ----

int uh_batch_create(void** handle, void* stream) {
    if (as_batch(stream)) return fail(UH_ERR_ARG, "uh_batch_create: stream is a batch");
    // device queries used by the persistent launchers, answered here and not during a capture
    (void)cu_count();
    (void)filter_wg_per_cu();
    (void)gemm_wg_per_cu();
    if (int rc = pagerank_binned_prepare()) return rc;
    Batch* b = new Batch();
    b->stream = (hipStream_t)stream;
    b->slot_bytes = 0;
    *handle = (void*)((uintptr_t)b | 1u);
    return UH_OK;
}

int uh_batch_destroy(void* handle) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_destroy: not a batch");
    hipStreamSynchronize(b->stream);
    for (int i = 0; i < Batch::RING; ++i) {
        for (auto& kv : b->graphs[i]) hipGraphExecDestroy(kv.second);
        b->graphs[i].clear();
        if (b->host_slot[i]) hipHostFree(b->host_slot[i]);
        if (b->dev_slot[i]) hipFree(b->dev_slot[i]);
        if (b->slot_free[i]) hipEventDestroy(b->slot_free[i]);
    }
    if (b->pr_scratch) hipFree(b->pr_scratch);
    delete b;
    return UH_OK;
}

int uh_is_batch(void* stream) { return as_batch(stream) ? 1 : 0; }

int uh_batch_pending(void* handle, int* n) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_pending: not a batch");
    size_t t = 0;
    for (int f = 0; f < FAM_COUNT; ++f) t += b->tiles[f].size();
    *n = (int)t;
    return UH_OK;
}

static const size_t kItemBytes[FAM_COUNT] = {
    sizeof(FilterItem), sizeof(uh_gemm_args), sizeof(uh_pagerank_args), sizeof(TrsmItem),
    sizeof(uh_pagerank_args), sizeof(uh_pagerank_binned_args)};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// The queue blocks of a flush move from the pinned host slot to the device slot by a kernel on
// the batch stream. A DMA-engine copy ahead of the persistent kernels made the kernels wait for
// the other engine: 12-43 us from the copy's end to the kernel's start in a kernel trace, where
// one kernel after another on the same queue starts within a few microseconds. Family f's block
// is words [off[f], off[f] + words[f]) of both slots; the loads are system-scope atomics, so no
// cache serves an old line of a host slot that the ring has reused.
struct UploadRanges {
    uint32_t off[FAM_COUNT];
    uint32_t words[FAM_COUNT];
};

__global__ void __launch_bounds__(256) k_upload_queue(uint32_t* dst, const uint32_t* src, UploadRanges r) {
    const uint32_t f = blockIdx.y;
    const uint32_t n = r.words[f];
    const uint32_t* s = src + r.off[f];
    uint32_t* d = dst + r.off[f];
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u)
        d[i] = __hip_atomic_load(s + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Algorithmic work of a flush that launches one family: GEMM items issue 2 M N K flop per
// product (times 3 for the split-BF16 products), filter items move 2 * rows * cols * 3 bytes
// (input and output once); other families report 0 (their callers know their work).
static double flush_work(const Batch* b, int mask) {
    if (__builtin_popcount(mask) != 1) return 0.0;
    double w = 0.0;
    if (mask == 1 << FAM_GEMM) {
        const uh_gemm_args* g = (const uh_gemm_args*)b->items[FAM_GEMM].data();
        for (size_t i = 0; i < b->tiles[FAM_GEMM].size(); ++i)
            w += 2.0 * g[i].M * g[i].N * g[i].K * ((g[i].precision & 0xff) == 0 ? 3.0 : 1.0);   // bits 8+: tile
    } else if (mask == 1 << FAM_FILTER) {
        const FilterItem* f = (const FilterItem*)b->items[FAM_FILTER].data();
        for (size_t i = 0; i < b->tiles[FAM_FILTER].size(); ++i) w += 6.0 * f[i].a.rows * f[i].a.cols;
    }
    return w;
}

// Layout of one upload: per family a QueueHeader block at a fixed offset of the slot
// (family f's block begins at f * family_stride) so that captured graphs can be replayed
// for any batch contents that fit the slot.
int uh_batch_flush(void* handle, int use_graph) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_flush: not a batch");
    size_t need_family[FAM_COUNT];
    size_t family_stride = 0;
    int mask = 0;
    for (int f = 0; f < FAM_COUNT; ++f) {
        size_t n = b->tiles[f].size();
        need_family[f] = align_up(sizeof(QueueHeader) + n * kItemBytes[f] + (n + 1) * 4, 256);
        if (need_family[f] > family_stride) family_stride = need_family[f];
        if (n) mask |= 1 << f;
    }
    if (!mask) return UH_OK;
    // Slots are sized generously and fixed per batch so graphs stay valid.
    size_t min_stride = 64 * 1024;
    if (family_stride < min_stride) family_stride = min_stride;
    size_t slot = family_stride * FAM_COUNT;
    if (slot > b->slot_bytes) {
        hipStreamSynchronize(b->stream);
        for (int i = 0; i < Batch::RING; ++i) {
            for (auto& kv : b->graphs[i]) hipGraphExecDestroy(kv.second);
            b->graphs[i].clear();
            if (b->host_slot[i]) hipHostFree(b->host_slot[i]);
            if (b->dev_slot[i]) hipFree(b->dev_slot[i]);
            b->host_slot[i] = nullptr; b->dev_slot[i] = nullptr;
            UH_HIP(hipHostMalloc((void**)&b->host_slot[i], slot, hipHostMallocDefault));
            UH_HIP(hipMalloc((void**)&b->dev_slot[i], slot));
            if (!b->slot_free[i]) UH_HIP(hipEventCreateWithFlags(&b->slot_free[i], hipEventDisableTiming));
        }
        b->slot_bytes = slot;
    }
    family_stride = b->slot_bytes / FAM_COUNT;
    if ((mask >> FAM_PRBLK) & 1) {
        size_t groups = 0;
        for (uint32_t t : b->tiles[FAM_PRBLK]) groups += t;
        const size_t need = groups * kPrGroupScratch;
        if (need > b->pr_scratch_bytes) {
            // earlier flushes on the stream may still read the old buffer
            UH_HIP(hipStreamSynchronize(b->stream));
            if (b->pr_scratch) UH_HIP(hipFree(b->pr_scratch));
            b->pr_scratch = nullptr;
            b->pr_scratch_bytes = 0;
            UH_HIP(hipMalloc(&b->pr_scratch, need));
            b->pr_scratch_bytes = need;
        }
    }
    int pos = b->ring_pos;
    b->ring_pos = (pos + 1) % Batch::RING;
    UH_HIP(hipEventSynchronize(b->slot_free[pos]));   // previous use of this slot has finished
    uint8_t* host = b->host_slot[pos];
    uint8_t* dev = b->dev_slot[pos];
    for (int f = 0; f < FAM_COUNT; ++f) {
        uint8_t* blk = host + f * family_stride;
        QueueHeader* h = (QueueHeader*)blk;
        uint32_t n = (uint32_t)b->tiles[f].size();
        h->next_tile = 0;
        h->n_items = n;
        h->items_offset = (uint32_t)sizeof(QueueHeader);
        h->prefix_offset = (uint32_t)align_up(sizeof(QueueHeader) + n * kItemBytes[f], 4);
        h->done = 0;
        memset(h->next_blk, 0, sizeof(h->next_blk));
        h->scratch = f == FAM_PRBLK ? (uint64_t)(uintptr_t)b->pr_scratch : 0;
        if (n) memcpy(blk + h->items_offset, b->items[f].data(), n * kItemBytes[f]);
        uint32_t* prefix = (uint32_t*)(blk + h->prefix_offset);
        uint32_t acc = 0;
        for (uint32_t i = 0; i < n; ++i) { prefix[i] = acc; acc += b->tiles[f][i]; }
        prefix[n] = acc;
        h->total_tiles = acc;
        g_items += n;
    }
    // The upload (up to the end of the last family present) goes on the stream ahead of the
    // kernels; the graphs hold only the kernel launches, so one capture per family mix serves
    // every batch, and the timing events bracket the kernels alone.
    UploadRanges ur;
    uint32_t max_words = 0;
    for (int f = 0; f < FAM_COUNT; ++f) {
        ur.off[f] = (uint32_t)(f * family_stride / 4);
        ur.words[f] = ((mask >> f) & 1) ? (uint32_t)(need_family[f] / 4) : 0u;   // need_family: 256-byte multiple
        if (ur.words[f] > max_words) max_words = ur.words[f];
    }
    const uint32_t ugx = std::min<uint32_t>((max_words + 255) / 256, 64u);
    UH_LAUNCH(k_upload_queue, dim3(ugx, FAM_COUNT), dim3(256), 0, b->stream, (uint32_t*)dev, (const uint32_t*)host, ur);
    UH_LAUNCH_CHECK("k_upload_queue");
    auto enqueue = [&](hipStream_t s) -> int {
        int rc = UH_OK;
        if ((mask >> FAM_FILTER) & 1) rc |= filter_persistent(dev + FAM_FILTER * family_stride, s, b->kinds[FAM_FILTER]);
        if ((mask >> FAM_GEMM) & 1) rc |= gemm_persistent(dev + FAM_GEMM * family_stride, s, b->kinds[FAM_GEMM]);
        if ((mask >> FAM_PAGERANK) & 1) rc |= pagerank_persistent(dev + FAM_PAGERANK * family_stride, s);
        if ((mask >> FAM_TRSM) & 1) rc |= trsm_persistent(dev + FAM_TRSM * family_stride, s);
        if ((mask >> FAM_PRBLK) & 1) rc |= pagerank_blocked_persistent(dev + FAM_PRBLK * family_stride, s);
        if ((mask >> FAM_PRBIN) & 1) rc |= pagerank_binned_persistent(dev + FAM_PRBIN * family_stride, s);
        return rc;
    };
    int rc;
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    std::unique_lock<std::mutex> tlock(b->tmtx);
    if (b->timing && b->timing_seq++ % b->timing == 0) {
        if (b->spare.empty()) {
            UH_HIP(hipEventCreate(&ev.first));
            UH_HIP(hipEventCreate(&ev.second));
        } else {
            ev = b->spare.back();
            b->spare.pop_back();
        }
    }
    // A graph pays off only when it replaces several launches: replaying a one-kernel graph costs
    // more host time (~25 us measured) than the direct launch it stands for (~6 us).
    if (use_graph && __builtin_popcount(mask) > 1) {
        const uint32_t key = (uint32_t)mask |
            (((mask >> FAM_FILTER) & 1) ? (uint32_t)filter_variant_for(b->kinds[FAM_FILTER]) << 8 : 0u) |
            (((mask >> FAM_GEMM) & 1) ? (uint32_t)gemm_variant_for(b->kinds[FAM_GEMM]) << 16 : 0u);
        hipGraphExec_t& ge = b->graphs[pos][key];
        if (!ge) {
            UH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue(b->stream);
            hipGraph_t g;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
        }
        if (ev.first) UH_HIP(hipEventRecord(ev.first, b->stream));
        UH_HIP(hipGraphLaunch(ge, b->stream));
        rc = UH_OK;
    } else {
        if (ev.first) UH_HIP(hipEventRecord(ev.first, b->stream));
        rc = enqueue(b->stream);
    }
    if (ev.first) {
        UH_HIP(hipEventRecord(ev.second, b->stream));
        b->timed.push_back({ev.first, ev.second, (uint32_t)mask, flush_work(b, mask)});
    }
    UH_HIP(hipEventRecord(b->slot_free[pos], b->stream));
    for (int f = 0; f < FAM_COUNT; ++f) { b->items[f].clear(); b->tiles[f].clear(); b->kinds[f] = 0; }
    return rc;
}

// Moves the finished timed flushes into the totals (caller holds tmtx).
static int drain_timed(Batch* b) {
    for (auto& t : b->timed) {
        UH_HIP(hipEventSynchronize(t.stop));
        float ms = 0.f;
        UH_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        b->timed_ms += ms;
        b->timed_flushes++;
        if (__builtin_popcount(t.mask) == 1) {
            const int f = __builtin_ctz(t.mask);
            b->fam_ms[f] += ms;
            b->fam_work[f] += t.work;
            b->fam_flushes[f]++;
        }
        b->spare.push_back({t.start, t.stop});
    }
    b->timed.clear();
    return UH_OK;
}

int uh_batch_set_timing(void* handle, int on) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_set_timing: not a batch");
    std::lock_guard<std::mutex> tlock(b->tmtx);
    b->timing = on > 0 ? (uint32_t)on : 0u;
    b->timing_seq = 0;
    return UH_OK;
}

int uh_batch_timing(void* handle, double* total_ms, uint64_t* flushes, int reset) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_timing: not a batch");
    std::lock_guard<std::mutex> tlock(b->tmtx);
    if (int rc = drain_timed(b)) return rc;
    if (total_ms) *total_ms = b->timed_ms;
    if (flushes) *flushes = b->timed_flushes;
    if (reset) {
        b->timed_ms = 0;
        b->timed_flushes = 0;
        for (int f = 0; f < FAM_COUNT; ++f) { b->fam_ms[f] = 0; b->fam_work[f] = 0; b->fam_flushes[f] = 0; }
    }
    return UH_OK;
}

int uh_batch_timing_family(void* handle, int family, double* ms, uint64_t* flushes, double* work, int reset) {
    Batch* b = as_batch(handle);
    if (!b) return fail(UH_ERR_ARG, "uh_batch_timing_family: not a batch");
    if (family < 0 || family >= FAM_COUNT) return fail(UH_ERR_ARG, "uh_batch_timing_family: no family %d", family);
    std::lock_guard<std::mutex> tlock(b->tmtx);
    if (int rc = drain_timed(b)) return rc;
    if (ms) *ms = b->fam_ms[family];
    if (flushes) *flushes = b->fam_flushes[family];
    if (work) *work = b->fam_work[family];
    if (reset) { b->fam_ms[family] = 0; b->fam_work[family] = 0; b->fam_flushes[family] = 0; }
    return UH_OK;
}

int uh_launch_counters(uint64_t* plain, uint64_t* persistent, uint64_t* items) {
    if (plain) *plain = g_plain_launches.load();
    if (persistent) *persistent = g_persistent_launches.load();
    if (items) *items = g_items.load();
    return UH_OK;
}

}  // extern "C"
