// pageRank kernels for gfx950 (replace pageRank_cuda / zeroInit,
// testSuite/pageRank/source/code/src/pageRankCUDA.cu:46-164).
//
// Two pulls over the HBM-resident in-link graph: the source-blocked one (below the CSR one),
// used whenever the graph's layout allows it, and the CSR-stream one described here.
//
// The reference pushes contributions with float atomics from every source page into a
// full-length output vector per subtask, over a padded record format copied from pageable host
// memory each iteration. Here the graph is resident in HBM as an in-link CSR (rows = destination
// pages, sources ascending) and rows are reduced CSR-stream style: a workgroup takes 256
// consecutive rows, streams their in-links through LDS in chunks of 2048 (each lane loads 8
// consecutive-in-CSR link ids, coalesced, then issues the 8 independent gathers of w), and each
// lane then sums its own row from LDS in index order. No atomics, a fixed summation order
// (deterministic), one pass over the edges per iteration, and many gathers in flight. The
// epilogue also produces the next iteration's per-source increment w = (damping * y) / deg, so
// the division of the reference's pageRank.cpp:262,367 happens once per page instead of once
// per edge.
#include "uh_internal.h"

namespace uh {

static constexpr int PR_THREADS = 256;
static constexpr int PR_ROWS_PER_TILE = PR_THREADS;          // one destination row per lane
static constexpr int PR_GATHER = 8;                          // gathers in flight per lane
static constexpr int PR_CHUNK = PR_THREADS * PR_GATHER;      // links staged in LDS per chunk

struct PrShared {
    uint32_t rp[PR_ROWS_PER_TILE + 1];      // row pointers of the tile's rows
    float val[PR_CHUNK];                    // w[col[k]] of the current chunk
};

__device__ __forceinline__ void pagerank_rows(const uh_pagerank_args& a, uint32_t tile, PrShared& sm) {
    const uint32_t tid = threadIdx.x;
    const uint32_t r0 = tile * PR_ROWS_PER_TILE;                 // first row (item-relative)
    const uint32_t nrows = min((uint32_t)PR_ROWS_PER_TILE, a.rows - r0);
    const uint32_t t0 = a.row0 + r0;                             // first row (global page id)
    const UH_GLOBAL uint32_t* row_ptr = as_global(a.row_ptr);
    const UH_GLOBAL uint32_t* col = as_global(a.col);
    const UH_GLOBAL float* w = as_global(a.w);
    for (uint32_t i = tid; i <= nrows; i += PR_THREADS) sm.rp[i] = row_ptr[t0 + i];
    __syncthreads();
    const uint32_t b = __builtin_amdgcn_readfirstlane(sm.rp[0]);
    const uint32_t e = __builtin_amdgcn_readfirstlane(sm.rp[nrows]);
    const uint32_t my_b = tid < nrows ? sm.rp[tid] : 0u;
    const uint32_t my_e = tid < nrows ? sm.rp[tid + 1] : 0u;
    float s = 0.f;
    for (uint32_t c0 = b; c0 < e; c0 += PR_CHUNK) {
        uint32_t src[PR_GATHER];
#pragma unroll
        for (int j = 0; j < PR_GATHER; ++j) {
            const uint32_t k = c0 + tid + j * PR_THREADS;
            src[j] = k < e ? col[k] : 0xffffffffu;
        }
#pragma unroll
        for (int j = 0; j < PR_GATHER; ++j) sm.val[tid + j * PR_THREADS] = src[j] != 0xffffffffu ? w[src[j]] : 0.f;
        __syncthreads();
        // this lane's row, restricted to the chunk, in link order
        const uint32_t lo = my_b > c0 ? my_b : c0;
        const uint32_t hi = my_e < c0 + PR_CHUNK ? my_e : c0 + PR_CHUNK;
        for (uint32_t k = lo; k < hi; ++k) s += sm.val[k - c0];
        __syncthreads();          // the chunk buffer is refilled next
    }
    if (tid < nrows) {
        const uint32_t t = t0 + tid;
        as_global(a.y)[t] = s;
        const uint32_t d = as_global(a.deg)[t];
        as_global(a.w_next)[t] = d ? (a.damping * s) / (float)d : 0.f;
    }
}

__global__ void __launch_bounds__(PR_THREADS) k_pagerank_pull(uh_pagerank_args a) {
    __shared__ PrShared sm;
    pagerank_rows(a, blockIdx.x, sm);
}

__global__ void __launch_bounds__(PR_THREADS) k_pagerank_persistent(const uint8_t* queue) {
    __shared__ PrShared sm;
    __shared__ uint32_t s_tile;
    QueueHeader* h = (QueueHeader*)queue;
    const uh_pagerank_args* items = (const uh_pagerank_args*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&h->next_tile, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t >= total) break;
        uint32_t item, local;
        queue_locate(prefix, n_items, t, item, local);
        pagerank_rows(uniform_load(items + item), local, sm);
        __syncthreads();          // LDS is reused by the next tile
    }
}

// ---- source-blocked pull ----
// The CSR pull above gathers w[src] for random sources over the whole vector (40 MB at 10 M
// pages), far more than one XCD's 4 MB L2, so nearly every gather is a miss served from the
// fabric. Over the source-blocked layout (see uh_pagerank_args) a unit of work is (row group g,
// source block x): the group's links from block x form one contiguous col segment, and all the
// gathers of the unit fall into w's block x (1/8 of the vector). The persistent kernel lets the
// workgroups of each XCD claim the units of "their" block first, so each XCD's L2 holds one block
// of w. A unit writes the partial sums of its group's 2048 rows; a second kernel adds the 8
// partials of each row in block order and produces y and w_next.
static constexpr uint32_t PR_GR = UH_PR_GROUP_ROWS;
static constexpr uint32_t PR_NB = UH_PR_BLOCKS;
static constexpr int PR_BCH = 4096;                          // links staged per chunk
static constexpr int PR_RPL = PR_GR / PR_THREADS;            // consecutive rows per lane
static_assert(PR_RPL == 8 && PR_NB == 8, "a lane owns 8 rows; 8 source blocks");
static constexpr int PR_NT = 2;                              // buffer cache policy: non-temporal

struct PrBlkShared {
    float val[PR_BCH];                      // w[col[k]] of the current chunk
    uint32_t wsum[PR_THREADS / 64];         // link totals of each wave's rows
};

// Unit (group g, block x): part[r] = sum of w[src] over the block-x links of row g*PR_GR + r,
// in link order, for r in [0, PR_GR). Lane t owns rows 8t .. 8t+7, whose links are consecutive.
__device__ __forceinline__ void pagerank_unit(const uh_pagerank_args& a, uint32_t g, uint32_t x,
                                              float* part_, PrBlkShared& sm) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t u = g * PR_NB + x;
    // The counts and link ids are read once: non-temporal loads (and stores of the partial
    // sums) leave the L2 to this XCD's block of w, which the gathers reuse.
    const __amdgpu_buffer_rsrc_t cnt_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)(a.bcnt + (size_t)u * PR_GR), 0, PR_GR, 0x00020000);
    const uint32_t c03 = __builtin_amdgcn_raw_buffer_load_b32(cnt_rsrc, 8 * tid, 0, PR_NT);
    const uint32_t c47 = __builtin_amdgcn_raw_buffer_load_b32(cnt_rsrc, 8 * tid + 4, 0, PR_NT);
    // row j's count, unpacked where it is used (keeps 6 registers free)
    auto cnt = [&](int j) { return ((j < 4 ? c03 : c47) >> (8 * (j & 3))) & 255u; };
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < PR_RPL; ++j) tot += cnt(j);
    uint32_t inc = tot;                      // inclusive scan of the lane totals in the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += v;
    }
    if (lane == 63) sm.wsum[wave] = inc;
    const UH_GLOBAL uint32_t* bseg = as_global(a.bseg);
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(bseg[u]);
    const uint32_t s1 = __builtin_amdgcn_readfirstlane(bseg[u + 1]);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < wave; ++k) base += sm.wsum[k];
    const uint32_t first = s0 + base + inc - tot;     // this lane's first link
    // Buffer loads: 32-bit offsets from a scalar base, so a gather costs one address register
    // (source ids are below kPrBlockedMaxPages = 2^30, which the layout builder enforces). The
    // col resource ends at the segment end, where loads return source 0; the val entries past
    // the segment are never summed.
    const __amdgpu_buffer_rsrc_t w_rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.w, 0, -1, 0x00020000);
    float acc[PR_RPL];
#pragma unroll
    for (int j = 0; j < PR_RPL; ++j) acc[j] = 0.f;
    for (uint32_t c0 = s0; c0 < s1; c0 += PR_BCH) {
        const uint32_t len = s1 - c0 < (uint32_t)PR_BCH ? s1 - c0 : (uint32_t)PR_BCH;
        const __amdgpu_buffer_rsrc_t c_rsrc =
            __builtin_amdgcn_make_buffer_rsrc((void*)(a.col + c0), 0, (int)(len * 4), 0x00020000);
        uint32_t src[PR_BCH / PR_THREADS];
#pragma unroll
        for (int j = 0; j < PR_BCH / PR_THREADS; ++j)
            src[j] = __builtin_amdgcn_raw_buffer_load_b32(c_rsrc, tid * 4u, j * PR_THREADS * 4, PR_NT);
#pragma unroll
        for (int j = 0; j < PR_BCH / PR_THREADS; ++j)
            sm.val[tid + j * PR_THREADS] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(w_rsrc, src[j] * 4u, 0, 0));
        __syncthreads();
        uint32_t k = first;
#pragma unroll
        for (int j = 0; j < PR_RPL; ++j) {
            const uint32_t e = k + cnt(j);
            const uint32_t lo = k > c0 ? k : c0;
            const uint32_t hi = e < c0 + PR_BCH ? e : c0 + PR_BCH;
            for (uint32_t q = lo; q < hi; ++q) acc[j] += sm.val[q - c0];
            k = e;
        }
        __syncthreads();          // the chunk buffer is refilled next
    }
    UH_GLOBAL float* part = as_global(part_ + (size_t)PR_RPL * tid);
#pragma unroll
    for (int j = 0; j < PR_RPL; ++j) __builtin_nontemporal_store(acc[j], part + j);
    __syncthreads();              // wsum is rewritten by the next unit
}

// Rows of group g inside the item: y = sum of the 8 block partials in block order, and w_next.
// part points at the group's partials (block x at part + x * PR_GR).
__device__ __forceinline__ void pagerank_combine(const uh_pagerank_args& a, uint32_t g, const float* part_) {
    const UH_GLOBAL float* part = as_global(part_);
    const UH_GLOBAL uint32_t* deg = as_global(a.deg);
    UH_GLOBAL float* y = as_global(a.y);
    UH_GLOBAL float* wn = as_global(a.w_next);
    const uint32_t end = a.row0 + a.rows;
#pragma unroll
    for (int j = 0; j < PR_RPL; ++j) {
        const uint32_t rl = threadIdx.x + j * PR_THREADS, r = g * PR_GR + rl;
        float t = 0.f;
#pragma unroll
        for (uint32_t x = 0; x < PR_NB; ++x) t += part[x * PR_GR + rl];
        if (r >= a.row0 && r < end) {
            y[r] = t;
            const uint32_t d = deg[r];
            wn[r] = d ? (a.damping * t) / (float)d : 0.f;
        }
    }
}

// Plain launch of one item: block b = unit (group row0 / PR_GR + b / 8, block b % 8). Workgroups
// are dispatched round-robin over the 8 XCDs, so block x runs on XCD x.
__global__ void __launch_bounds__(PR_THREADS) k_pagerank_blocked(uh_pagerank_args a, float* scratch) {
    __shared__ PrBlkShared sm;
    const uint32_t b = blockIdx.x;
    pagerank_unit(a, a.row0 / PR_GR + b / PR_NB, b % PR_NB, scratch + (size_t)b * PR_GR, sm);
}

__global__ void __launch_bounds__(PR_THREADS) k_pagerank_combine(uh_pagerank_args a, const float* scratch) {
    const uint32_t gl = blockIdx.x;
    pagerank_combine(a, a.row0 / PR_GR + gl, scratch + (size_t)gl * PR_NB * PR_GR);
}

// Batched: tiles of the queue are (item, local group) pairs, k = the global group index; unit
// (k, x) writes its partials to scratch + (k * 8 + x) * PR_GR. A workgroup claims the units of
// its own XCD's block first, then helps with the other blocks, so every unit is taken whatever
// the placement of the workgroups. Held to 64 VGPRs (8 waves per SIMD, all 8 workgroups of a CU
// resident), as the plain kernel is by itself: the gathers need every wave in flight.
__global__ void __launch_bounds__(PR_THREADS)
k_pagerank_blocked_persistent(const uint8_t* queue) {
    __shared__ PrBlkShared sm;
    __shared__ uint32_t s_k;
    QueueHeader* h = (QueueHeader*)queue;
    const uh_pagerank_args* items = (const uh_pagerank_args*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    float* scratch = (float*)h->scratch;
    const uint32_t x0 = __builtin_amdgcn_readfirstlane(xcc_id());
    for (uint32_t i = 0; i < PR_NB; ++i) {
        const uint32_t x = (x0 + i) % PR_NB;
        for (;;) {
            if (threadIdx.x == 0) s_k = atomicAdd(&h->next_blk[x * 32], 1u);
            __syncthreads();
            const uint32_t k = __builtin_amdgcn_readfirstlane(s_k);
            __syncthreads();
            if (k >= total) break;
            uint32_t item, local;
            queue_locate(prefix, n_items, k, item, local);
            const uh_pagerank_args a = uniform_load(items + item);
            pagerank_unit(a, a.row0 / PR_GR + local, x, scratch + ((size_t)k * PR_NB + x) * PR_GR, sm);
        }
    }
}

__global__ void __launch_bounds__(PR_THREADS) k_pagerank_combine_persistent(const uint8_t* queue) {
    const QueueHeader* h = (const QueueHeader*)queue;
    const uh_pagerank_args* items = (const uh_pagerank_args*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    const float* scratch = (const float*)h->scratch;
    for (uint32_t k = blockIdx.x; k < total; k += gridDim.x) {
        uint32_t item, local;
        queue_locate(prefix, n_items, k, item, local);
        const uh_pagerank_args a = uniform_load(items + item);
        pagerank_combine(a, a.row0 / PR_GR + local, scratch + (size_t)k * PR_NB * PR_GR);
    }
}

__global__ void __launch_bounds__(256) k_pagerank_init_w(float* w, const uint32_t* deg, uint32_t n,
                                                         float x0, float damping) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) {
        uint32_t d = deg[i];
        w[i] = d ? (damping * x0) / (float)d : 0.f;
    }
}

// Reference-style push over padded records (count, t_1..t_count, 0-padding to max_out); targets
// are 1-based page ids (pageRankCUDA.cu:46-65). Float atomics: summation order is not fixed.
__global__ void __launch_bounds__(256) k_pagerank_push_ell(const uint32_t* rec, uint32_t pages,
                                                           uint32_t max_out, const float* x, float x0,
                                                           int first, float damping, float* y) {
    uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= pages) return;
    const uint32_t* r = rec + (uint64_t)p * (max_out + 1);
    const uint32_t n = r[0];
    const float incr = (damping * (first ? x0 : x[p])) / (float)n;
    for (uint32_t k = 0; k < n; ++k) atomicAdd(&y[r[1 + k] - 1], incr);
}

uint32_t pagerank_tiles(const uh_pagerank_args& a) {
    return (a.rows + PR_ROWS_PER_TILE - 1) / PR_ROWS_PER_TILE;
}

uint32_t pagerank_groups(const uh_pagerank_args& a) {
    return (a.row0 + a.rows - 1) / PR_GR - a.row0 / PR_GR + 1;
}

int pagerank_blocked_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_pagerank_blocked_persistent, dim3(cu_count() * 8), dim3(PR_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_pagerank_blocked_persistent");
    UH_LAUNCH(k_pagerank_combine_persistent, dim3(cu_count() * 8), dim3(PR_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_pagerank_combine_persistent");
    g_persistent_launches += 2;
    return UH_OK;
}

int pagerank_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_pagerank_persistent, dim3(cu_count() * 8), dim3(PR_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_pagerank_persistent");
    g_persistent_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

// Partial sums of plain source-blocked launches, per stream: launches on one stream run in
// stream order and share the buffer.
static std::map<hipStream_t, std::pair<float*, size_t>> g_pr_scratch;
static std::mutex g_pr_mtx;

extern "C" int uh_pagerank_pull(const uh_pagerank_args* a, void* stream) {
    if (!a || !a->row_ptr || !a->w || !a->y || !a->w_next || !a->deg)
        return fail(UH_ERR_ARG, "uh_pagerank_pull: null pointer");
    if (a->rows == 0) return UH_OK;
    if (!a->col) return fail(UH_ERR_ARG, "uh_pagerank_pull: null col");
    if (a->bcnt) {
        if (!a->bseg) return fail(UH_ERR_ARG, "uh_pagerank_pull: bcnt without bseg");
        if ((uintptr_t)a->bcnt & 3u) return fail(UH_ERR_ARG, "uh_pagerank_pull: bcnt not 4-byte aligned");
        if ((uint64_t)a->row0 + a->rows > kPrBlockedMaxPages)
            return fail(UH_ERR_ARG, "uh_pagerank_pull: the source-blocked layout holds pages below 2^30");
        const uint32_t groups = pagerank_groups(*a);
        if (Batch* b = as_batch(stream)) {
            record_item(b, FAM_PRBLK, a, sizeof(*a), groups);
            return UH_OK;
        }
        hipStream_t s = (hipStream_t)stream;
        const size_t need = groups * kPrGroupScratch;
        float* scratch;
        {
            std::lock_guard<std::mutex> lk(g_pr_mtx);
            auto& e = g_pr_scratch[s];
            if (e.second < need) {
                if (e.first) { UH_HIP(hipStreamSynchronize(s)); UH_HIP(hipFree(e.first)); e.first = nullptr; e.second = 0; }
                UH_HIP(hipMalloc(&e.first, need));
                e.second = need;
            }
            scratch = e.first;
        }
        UH_LAUNCH(k_pagerank_blocked, dim3(groups * PR_NB), dim3(PR_THREADS), 0, s, *a, scratch);
        UH_LAUNCH_CHECK("k_pagerank_blocked");
        UH_LAUNCH(k_pagerank_combine, dim3(groups), dim3(PR_THREADS), 0, s, *a, scratch);
        UH_LAUNCH_CHECK("k_pagerank_combine");
        g_plain_launches += 2;
        return UH_OK;
    }
    const uint32_t tiles = pagerank_tiles(*a);
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_PAGERANK, a, sizeof(*a), tiles);
        return UH_OK;
    }
    UH_LAUNCH(k_pagerank_pull, dim3(tiles), dim3(PR_THREADS), 0, (hipStream_t)stream, *a);
    UH_LAUNCH_CHECK("k_pagerank_pull");
    g_plain_launches++;
    return UH_OK;
}

// ---- allocation-free variant of uh_pagerank_pull ----
// The partial sums of the source-blocked kernels live in a static device array instead of a
// buffer allocated per stream. Each GPU the library runs on holds its own copy (32 MiB) until the
// process ends; host code reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in, not a device address).
static constexpr uint32_t kNoallocGroups = 512;              // row groups: 1,048,576 rows
static __device__ __align__(256) float pr_noalloc_partials[(size_t)kNoallocGroups * PR_NB * PR_GR];
static constexpr int kMaxDevices = 64;
static float* g_noalloc_ptr[kMaxDevices];
static std::mutex g_noalloc_mtx;

// Device address of the current device's copy, looked up once per device.
static int noalloc_partials(float** out) {
    int dev = 0;
    UH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mtx);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        UH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(pr_noalloc_partials)));
        if (!p) return fail(UH_ERR_HIP, "uh_pagerank_pull_noalloc: no device address for the partial-sum array");
        g_noalloc_ptr[dev] = (float*)p;
    }
    *out = g_noalloc_ptr[dev];
    return UH_OK;
}

extern "C" int uh_pagerank_pull_noalloc(const uh_pagerank_args* a, void* stream) {
    if (!a || !a->row_ptr || !a->w || !a->y || !a->w_next || !a->deg)
        return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: null pointer");
    if (as_batch(stream))
        return fail(UH_ERR_BATCH, "uh_pagerank_pull_noalloc: takes a stream, not a batch handle (a flush may allocate)");
    if (a->rows == 0) return UH_OK;
    if (!a->col) return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: null col");
    hipStream_t s = (hipStream_t)stream;
    if (!a->bcnt) {                      // the CSR pull keeps no partial sums
        UH_LAUNCH(k_pagerank_pull, dim3(pagerank_tiles(*a)), dim3(PR_THREADS), 0, s, *a);
        UH_LAUNCH_CHECK("k_pagerank_pull");
        g_plain_launches++;
        return UH_OK;
    }
    if (!a->bseg) return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: bcnt without bseg");
    if ((uintptr_t)a->bcnt & 3u) return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: bcnt not 4-byte aligned");
    if ((uint64_t)a->row0 + a->rows > kPrBlockedMaxPages)
        return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: the source-blocked layout holds pages below 2^30");
    const uint32_t groups = pagerank_groups(*a);
    if (groups > kNoallocGroups)
        return fail(UH_ERR_ARG, "uh_pagerank_pull_noalloc: rows [%u, %u) span %u row groups of %u; at most %u fit "
                    "the static partial-sum array", a->row0, a->row0 + a->rows, groups, PR_GR, kNoallocGroups);
    float* part = nullptr;
    if (int rc = noalloc_partials(&part)) return rc;
    // every unit writes all PR_GR partials of its (group, block) before the combine reads them,
    // so the array needs no clearing between calls
    UH_LAUNCH(k_pagerank_blocked, dim3(groups * PR_NB), dim3(PR_THREADS), 0, s, *a, part);
    UH_LAUNCH_CHECK("k_pagerank_blocked");
    UH_LAUNCH(k_pagerank_combine, dim3(groups), dim3(PR_THREADS), 0, s, *a, part);
    UH_LAUNCH_CHECK("k_pagerank_combine");
    g_plain_launches += 2;
    return UH_OK;
}

extern "C" int uh_pagerank_init_w(float* w, const uint32_t* deg, uint32_t n, float x0, float damping,
                                  void* stream) {
    if (!n) return UH_OK;
    if (!w || !deg) return fail(UH_ERR_ARG, "uh_pagerank_init_w: null pointer");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_pagerank_init_w, dim3((n + 255) / 256), dim3(256), 0, s, w, deg, n, x0, damping);
    UH_LAUNCH_CHECK("k_pagerank_init_w");
    g_plain_launches++;
    return UH_OK;
}

extern "C" int uh_pagerank_push_ell(const uint32_t* records, uint32_t pages, uint32_t max_out,
                                    const float* x, float x0, int first_iteration, float damping,
                                    float* y, void* stream) {
    if (!pages) return UH_OK;
    if (!records || !y || (!first_iteration && !x)) return fail(UH_ERR_ARG, "uh_pagerank_push_ell: null pointer");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_pagerank_push_ell, dim3((pages + 255) / 256), dim3(256), 0, s,
                       records, pages, max_out, x, x0, first_iteration, damping, y);
    UH_LAUNCH_CHECK("k_pagerank_push_ell");
    g_plain_launches++;
    return UH_OK;
}
