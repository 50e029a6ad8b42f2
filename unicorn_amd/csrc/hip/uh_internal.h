// Internal declarations shared by the HIP sources of libunicorn_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <atomic>
#include <map>
#include <mutex>

#include "unicorn_hip.h"

namespace uh {

// Records the failure text for uh_last_error() and returns the status.
int fail(int status, const char* fmt, ...);
int check_hip(hipError_t err, const char* what);

#define UH_HIP(call)                                            \
    do {                                                        \
        hipError_t uh_e_ = (call);                              \
        if (uh_e_ != hipSuccess) return uh::check_hip(uh_e_, #call); \
    } while (0)

// Launch with a clean error slot, so that UH_LAUNCH_CHECK reports this launch only and not an
// earlier failure that was already returned to its caller.
#define UH_LAUNCH(...)                                          \
    do {                                                        \
        (void)hipGetLastError();                                \
        hipLaunchKernelGGL(__VA_ARGS__);                        \
    } while (0)

#define UH_LAUNCH_CHECK(name)                                   \
    do {                                                        \
        hipError_t uh_e_ = hipGetLastError();                   \
        if (uh_e_ != hipSuccess) return uh::check_hip(uh_e_, name); \
    } while (0)

extern std::atomic<uint64_t> g_plain_launches;
extern std::atomic<uint64_t> g_persistent_launches;
extern std::atomic<uint64_t> g_items;

int cu_count();

// Kernel families that a batch can hold. Each family records fixed-size descriptors.
// FAM_PRBLK holds pageRank items over the source-blocked layout; their tiles are row groups.
enum Family { FAM_FILTER = 0, FAM_GEMM = 1, FAM_PAGERANK = 2, FAM_TRSM = 3, FAM_PRBLK = 4, FAM_PRBIN = 5,
              FAM_COUNT = 6 };

// Work item of the filter family: the launcher's arguments (pad = view width in pixels) plus the
// dot4 weight vectors of radius <= 2, packed once on the host (wv[dy][q][d], see imagefilter.hip).
struct FilterWeights {
    uint32_t wv[5 * 4 * 2];
    int32_t wsum, pad[3];
};
struct FilterItem {
    uh_imagefilter_args a;
    FilterWeights w;
};

// Work item of the TRSM family (both triangular kinds of the LU stage 2).
struct TrsmItem {
    const float* tri; float* x;
    uint64_t ldt, ldx;
    int32_t n, m, kind, pad;     // kind 0 = L^-1 X (lower non-unit), 1 = X U^-1 (upper unit)
};

// Device work-queue block of one family, as uploaded for one flush:
//   QueueHeader | item descriptors (n_items * item_size) | tile prefix (n_items + 1 uint32)
struct QueueHeader {
    uint32_t next_tile;       // work-stealing counter, zeroed by the upload itself
    uint32_t n_items;
    uint32_t total_tiles;
    uint32_t items_offset;    // bytes from the header start
    uint32_t prefix_offset;   // bytes from the header start
    uint32_t done;            // workgroups finished (for the last-one-out reset)
    uint32_t pad[2];
    uint64_t scratch;         // device scratch of the flush (FAM_PRBLK partial sums), or 0
    uint32_t pad2[22];
    // Per source block claim counters of FAM_PRBLK (counter x at next_blk[x * 32]), zeroed by the
    // upload. Each sits in its own 128-byte line: on one shared line the claims of the 8 XCDs
    // contended and cost a fifth of the kernel's time.
    uint32_t next_blk[8 * 32];
};
static_assert(sizeof(QueueHeader) % 128 == 0 && offsetof(QueueHeader, next_blk) % 128 == 0,
              "claim counters on their own lines; work items follow the header");

struct Batch {
    hipStream_t stream = nullptr;
    std::vector<uint8_t> items[FAM_COUNT];     // raw descriptors per family
    std::vector<uint32_t> tiles[FAM_COUNT];    // tile count of each recorded item
    uint32_t kinds[FAM_COUNT] = {};            // OR of the kernel-variant bits of the items
    // upload ring: pinned host slot + device slot per ring entry
    static const int RING = 4;
    uint8_t* host_slot[RING] = {};
    uint8_t* dev_slot[RING] = {};
    hipEvent_t slot_free[RING] = {};
    size_t slot_bytes = 0;
    int ring_pos = 0;
    // graphs captured per ring slot, keyed by family mask and the filter kernel variant
    std::map<uint32_t, hipGraphExec_t> graphs[RING];
    // device time of the flushes (events recorded on the batch stream around each flush)
    uint32_t timing = 0;          // events around every timing-th flush (0: none)
    uint64_t timing_seq = 0;      // flushes since timing was set
    std::mutex tmtx;                 // guards timed / spare / the totals
    struct Timed {
        hipEvent_t start, stop;
        uint32_t mask;                   // families the flush launched
        double work;                     // algorithmic work of a one-family flush (see uh_batch_timing_family)
    };
    std::vector<Timed> timed;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
    double timed_ms = 0;
    uint64_t timed_flushes = 0;
    // flushes that launched one family only, per family
    double fam_ms[FAM_COUNT] = {}, fam_work[FAM_COUNT] = {};
    uint64_t fam_flushes[FAM_COUNT] = {};
    // partial sums of the FAM_PRBLK items of one flush (flushes on the batch stream run in
    // order, so one buffer serves all of them); grown when a flush needs more
    void* pr_scratch = nullptr;
    size_t pr_scratch_bytes = 0;
};

Batch* as_batch(void* stream);   // nullptr if the stream is not a batch handle
// Stream for an operation that cannot be recorded as a work item. A batch handle acts as a
// barrier: the items recorded so far are flushed, then the operation runs on the batch's stream.
hipStream_t plain_stream_of(void* stream, int* rc);
// kind_bits: bits of the kernel variants the item can run on (a family with one variant passes 1).
void record_item(Batch* b, Family f, const void* desc, size_t bytes, uint32_t tiles,
                 uint32_t kind_bits = 1);

// Per-family persistent launchers used by uh_batch_flush (queue = device pointer to the
// family's QueueHeader).
int filter_persistent(const uint8_t* queue, hipStream_t s, uint32_t kinds);
// Filter kernel variant that a flush with these kind bits launches.
int filter_variant_for(uint32_t kinds);
int gemm_persistent(const uint8_t* queue, hipStream_t s, uint32_t kinds);
// GEMM kernel variant that a flush with these kind bits launches.
int gemm_variant_for(uint32_t kinds);
// Workgroups of the persistent GEMM kernels resident per CU (occupancy queries, cached).
int gemm_wg_per_cu();
int pagerank_persistent(const uint8_t* queue, hipStream_t s);
int pagerank_blocked_persistent(const uint8_t* queue, hipStream_t s);
int trsm_persistent(const uint8_t* queue, hipStream_t s);
int pagerank_binned_persistent(const uint8_t* queue, hipStream_t s);
// Raises the binned pageRank kernels' LDS limit (once; outside any graph capture).
int pagerank_binned_prepare();
// Workgroups of the persistent filter kernel resident per CU (an occupancy query, cached).
int filter_wg_per_cu();

// Tile counts of one item of each family (host side, for the queue prefix).
uint32_t filter_tiles(const FilterItem& it);
uint32_t gemm_tiles(const uh_gemm_args& a);
uint32_t pagerank_tiles(const uh_pagerank_args& a);
uint32_t pagerank_groups(const uh_pagerank_args& a);   // row groups of a source-blocked item
// Bytes of partial sums that one row group of a source-blocked item needs.
constexpr size_t kPrGroupScratch = (size_t)UH_PR_BLOCKS * UH_PR_GROUP_ROWS * sizeof(float);
// Pages of a graph in the source-blocked layout stay below 2^30: the gathers address w with a
// 32-bit byte offset (source id * 4) from the vector's base.
constexpr uint64_t kPrBlockedMaxPages = 1ull << 30;
uint32_t trsm_tiles(const TrsmItem& a);

// Address spaces. A pointer that the compiler cannot trace to a kernel argument (for example
// one read from a work item) is "generic" and its accesses become FLAT instructions, which count
// against the LDS wait counter too: every LDS wait would then also wait for outstanding global
// loads and stores. Kernels access HBM through global-space pointers instead.
#define UH_GLOBAL __attribute__((address_space(1)))
#define UH_CONSTANT __attribute__((address_space(4)))

template <typename T>
__device__ __forceinline__ const UH_GLOBAL T* as_global(const T* p) { return (const UH_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ UH_GLOBAL T* as_global(T* p) { return (UH_GLOBAL T*)p; }

// Device helper: load a work item whose address is the same in every lane into scalar registers
// with scalar loads. Valid for memory that no lane writes while the kernel runs (work items are
// uploaded before the launch; the scalar cache is invalidated at each dispatch).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "work items are dword-sized");
    T v;
    const UH_CONSTANT uint32_t* s = (const UH_CONSTANT uint32_t*)p;
    uint32_t* d = (uint32_t*)&v;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = s[i];
    return v;
}

// Device helper: map a global tile index of a work queue to (item, local tile) by a binary
// search over the prefix array (prefix[i] = first tile of item i). With a uniform tile index the
// search runs on scalar loads (the prefix is read-only while the kernel runs).
__device__ inline void queue_locate(const uint32_t* prefix, uint32_t n_items, uint32_t tile,
                                    uint32_t& item, uint32_t& local) {
    const UH_CONSTANT uint32_t* p = (const UH_CONSTANT uint32_t*)prefix;
    uint32_t lo = 0, hi = n_items;           // invariant: prefix[lo] <= tile < prefix[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (p[mid] <= tile) lo = mid; else hi = mid;
    }
    item = lo;
    local = tile - p[lo];
}

// XCD of the calling wave: hardware register XCC_ID (id 20), bits [3:0]. Used only to choose
// which work a workgroup claims first; every claim loop also takes the other XCDs' work.
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & 7u;
}

// The same search over a copy of the prefix array in LDS (see stage_prefix).
__device__ inline void queue_locate_lds(const uint32_t* s_prefix, uint32_t n_items, uint32_t tile,
                                        uint32_t& item, uint32_t& local) {
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_prefix[mid] <= tile) lo = mid; else hi = mid;
    }
    item = __builtin_amdgcn_readfirstlane(lo);
    local = tile - __builtin_amdgcn_readfirstlane(s_prefix[item]);
}

// Copies prefix[0..n_items] into LDS (the caller synchronises before use); false when it does
// not fit in `cap` entries.
__device__ inline bool stage_prefix(const uint32_t* prefix, uint32_t n_items, uint32_t* s_prefix, uint32_t cap) {
    if (n_items + 1 > cap) return false;
    for (uint32_t i = threadIdx.x; i <= n_items; i += blockDim.x) s_prefix[i] = as_global(prefix)[i];
    return true;
}

}  // namespace uh
