// pageRankInvertedIndex kernels. The reference's GPU callbacks for this app are empty stubs
// (invertIndex_cudaLaunchFunc / pageRank_cudaLaunchFunc, pageRankInvertedIndexCUDA.cu:92-100);
// these kernels do on the device what its CPU callbacks do (pageRankInvertedIndex.cpp:301-382,
// 438-472), with the same results:
//
// uh_invert_index: one subtask's pages of the web dump (per page: out-link count c, then c 1-based
// target pages, padded to max_out) become inverted-index entries {src page, dest page (0-based),
// src out-links}, grouped by the subtask that owns the destination page (dest / P). Entries of
// group d occupy positions [S_d, S_d + n_d) of the output window, S_d = n_0 + ... + n_(d-1), and
// within a group run in REVERSE enumeration order (page order, then link order): the reference
// fills each group from its highest position down. n_d goes to totals[d].
//   gfx950 design. Pages are cut into tiles (at most 4096, a multiple of 256 pages each);
//   1. k_ii_count: per tile, a group histogram in LDS -> tile_count[d][tile];
//   2. k_ii_scan: per group, an exclusive scan over the tiles (tile_base) and the group total;
//      k_ii_starts: the exclusive scan over the groups (S_d);
//   3. k_ii_scatter: one wave per tile walks the tile's links in enumeration order, 64 at a time
//      (lanes spread over pages by a wave prefix sum of the counts); a link's rank among the
//      tile's earlier links of its group is the group's running count plus the lanes below it
//      with the same group (one ballot per distinct group in the step). Its position is
//      S_d + n_d - 1 - (tile_base[d][tile] + rank).
//   Records with c > max_out and links outside [1, total_pages] are skipped by all kernels alike
//   and reported as an error after the call (the call waits for its stream).
// uh_prii_rank: out[dest - first] = sum over the entries of damping * (x / out-links) with x the
// source's rank (or `initial`), accumulated with float atomics (summation order differs from the
// CPU callback's entry order: results agree within float rounding, not bit for bit).
#include "uh_internal.h"

#include <map>
#include <mutex>

namespace uh {

static constexpr uint32_t II_MAX_TILES = 4096;
static constexpr uint32_t II_MAX_GROUPS = 8192;     // LDS histogram / running counts (32 KB)

struct IiEntry {
    uint32_t src, dest, outlinks;
};

__device__ __forceinline__ bool ii_link_ok(uint32_t link, uint32_t total) { return link >= 1 && link <= total; }

__global__ void __launch_bounds__(256) k_ii_count(const uint32_t* dump, uint32_t pages, uint32_t max_out,
                                                   uint32_t total, uint32_t P, uint32_t groups, uint32_t tile_pages,
                                                   uint32_t tiles, uint32_t* tile_count, uint32_t* bad) {
    extern __shared__ uint32_t hist[];
    for (uint32_t d = threadIdx.x; d < groups; d += 256) hist[d] = 0;
    __syncthreads();
    const uint32_t p0 = blockIdx.x * tile_pages, p1 = min(pages, p0 + tile_pages);
    const uint64_t rec = (uint64_t)max_out + 1;
    uint32_t nbad = 0;
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += 256) {
        const uint32_t* r = dump + (uint64_t)p * rec;
        const uint32_t c = r[0];
        if (c > max_out) { ++nbad; continue; }
        for (uint32_t k = 0; k < c; ++k) {
            const uint32_t link = r[1 + k];
            if (!ii_link_ok(link, total)) { ++nbad; continue; }
            atomicAdd(&hist[(link - 1) / P], 1u);
        }
    }
    if (nbad) atomicAdd(bad, nbad);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < groups; d += 256) tile_count[(uint64_t)d * tiles + blockIdx.x] = hist[d];
}

// wave-wide inclusive scan (64 lanes)
__device__ __forceinline__ uint32_t ii_wave_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// One block per group: exclusive scan of tile_count[d][*] in place; totals[d] = the sum.
__global__ void __launch_bounds__(256) k_ii_scan(uint32_t* tile_count, uint32_t tiles, uint32_t* totals) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    uint32_t* row = tile_count + (uint64_t)blockIdx.x * tiles;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < tiles; t0 += 256) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < tiles ? row[t] : 0;
        const uint32_t inc = ii_wave_scan(v);
        if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
        __syncthreads();
        uint32_t before = carry + inc - v;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) before += wsum[w];
        if (t < tiles) row[t] = before;
        __syncthreads();
        if (threadIdx.x == 255) carry = before + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// One block: starts[d] = totals[0] + ... + totals[d-1].
__global__ void __launch_bounds__(256) k_ii_starts(const uint32_t* totals, uint32_t groups, uint32_t* starts) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t d0 = 0; d0 < groups; d0 += 256) {
        const uint32_t d = d0 + threadIdx.x;
        const uint32_t v = d < groups ? totals[d] : 0;
        const uint32_t inc = ii_wave_scan(v);
        if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
        __syncthreads();
        uint32_t before = carry + inc - v;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) before += wsum[w];
        if (d < groups) starts[d] = before;
        __syncthreads();
        if (threadIdx.x == 255) carry = before + v;
        __syncthreads();
    }
}

// One wave per tile. running[] (LDS, groups entries) counts the tile's links of each group so far.
__global__ void __launch_bounds__(64) k_ii_scatter(const uint32_t* dump, uint32_t pages, uint32_t max_out,
                                                  uint32_t total, uint32_t P, uint32_t groups, uint32_t first_page,
                                                  uint32_t tile_pages, uint32_t tiles, const uint32_t* tile_base,
                                                  const uint32_t* totals, const uint32_t* starts, IiEntry* out) {
    extern __shared__ uint32_t running[];
    __shared__ uint32_t offs[64];
    const uint32_t lane = threadIdx.x;
    for (uint32_t d = lane; d < groups; d += 64) running[d] = 0;
    __syncthreads();
    const uint32_t p0 = blockIdx.x * tile_pages, p1 = min(pages, p0 + tile_pages);
    const uint64_t rec = (uint64_t)max_out + 1;
    for (uint32_t g0 = p0; g0 < p1; g0 += 64) {
        // 64 pages: lane l holds page g0 + l's link count (0 for an invalid record)
        const uint32_t p = g0 + lane;
        uint32_t c = 0;
        if (p < p1) {
            c = dump[(uint64_t)p * rec];
            if (c > max_out) c = 0;
        }
        const uint32_t inc = ii_wave_scan(c);
        offs[lane] = inc - c;                       // first link slot of page g0 + lane
        const uint32_t links = __shfl(inc, 63, 64);
        __syncthreads();
        for (uint32_t s0 = 0; s0 < links; s0 += 64) {
            const uint32_t slot = s0 + lane;
            bool live = slot < links;
            uint32_t page_l = 0, link = 0, outl = 0, d = 0;
            if (live) {
                // last page whose first slot <= slot (pages with no links share offsets)
                uint32_t lo = 0, hi = 63;
                while (lo < hi) {
                    const uint32_t m = (lo + hi + 1) >> 1;
                    if (offs[m] <= slot) lo = m;
                    else hi = m - 1;
                }
                page_l = lo;
                const uint32_t* r = dump + (uint64_t)(g0 + page_l) * rec;
                outl = r[0];
                link = r[1 + (slot - offs[page_l])];
                live = ii_link_ok(link, total);
                if (live) d = (link - 1) / P;
            }
            // rank among the lower lanes of the same group, one distinct group per round
            uint64_t pending = __ballot(live);
            uint32_t rank = 0;
            while (pending) {
                const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
                const uint32_t d0 = __shfl(d, leader, 64);
                const uint64_t same = __ballot(live && d == d0) & pending;
                const uint32_t base = running[d0];
                if (live && d == d0) {
                    const uint64_t below = same & ((1ull << lane) - 1ull);
                    rank = base + (uint32_t)__popcll(below);
                }
                __syncthreads();
                if (lane == leader) running[d0] = base + (uint32_t)__popcll(same);
                __syncthreads();
                pending &= ~same;
            }
            if (live) {
                const uint32_t r = tile_base[(uint64_t)d * tiles + blockIdx.x] + rank;
                const uint32_t pos = starts[d] + totals[d] - 1u - r;
                IiEntry e;
                e.src = first_page + g0 + page_l;
                e.dest = link - 1;
                e.outlinks = outl;
                out[pos] = e;
            }
        }
        __syncthreads();                            // offs reuse
    }
}

__global__ void __launch_bounds__(256) k_prii_rank(const IiEntry* e, uint32_t n, const float* in, float initial,
                                                   float damping, uint32_t first, float* out, uint32_t out_pages,
                                                   uint32_t total) {
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += stride) {
        const IiEntry x = e[i];
        const uint32_t o = x.dest - first;
        if (o >= out_pages || x.src >= total || !x.outlinks) continue;
        const float r = in ? in[x.src] : initial;
        const float q = r / (float)x.outlinks;
        unsafeAtomicAdd(out + o, damping * q);
    }
}

__global__ void __launch_bounds__(256) k_prii_zero(float* out, uint32_t n) {
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += stride) out[i] = 0.f;
}

// Work buffers of uh_invert_index: one per (device, stream), grown on demand, kept while the
// process runs; layout: bad counter | starts[groups] | tile_count[groups][tiles].
static std::mutex g_ii_mtx;
static std::map<std::pair<int, hipStream_t>, std::pair<uint32_t*, size_t>> g_ii;

static uint32_t* ii_buffer(hipStream_t stream, size_t words) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(g_ii_mtx);
    auto& e = g_ii[std::make_pair(dev, stream)];
    if (e.second < words) {
        if (e.first && hipFree(e.first) != hipSuccess) return nullptr;
        e.first = nullptr;
        e.second = 0;
        if (hipMalloc((void**)&e.first, words * 4) != hipSuccess) return nullptr;
        e.second = words;
    }
    return e.first;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_invert_index(const uint32_t* dump, uint32_t pages, uint32_t max_out, uint32_t first_page,
                               uint32_t total_pages, uint32_t pages_per_subtask, uint32_t groups, void* out_entries,
                               uint32_t* totals, void* stream) {
    if (!groups || !pages_per_subtask) return fail(UH_ERR_ARG, "uh_invert_index: no groups");
    if (groups > II_MAX_GROUPS) return fail(UH_ERR_ARG, "uh_invert_index: %u groups (at most %u)", groups, II_MAX_GROUPS);
    if ((uint64_t)groups * pages_per_subtask < total_pages)
        return fail(UH_ERR_ARG, "uh_invert_index: %u groups of %u pages do not cover %u pages", groups,
                    pages_per_subtask, total_pages);
    if (!totals) return fail(UH_ERR_ARG, "uh_invert_index: null totals");
    if (pages && (!dump || !out_entries)) return fail(UH_ERR_ARG, "uh_invert_index: null buffer");
    if (((uintptr_t)dump | (uintptr_t)out_entries | (uintptr_t)totals) & 3u)
        return fail(UH_ERR_ARG, "uh_invert_index: buffers must be 4-byte aligned");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    uint32_t tile_pages = 256;
    while ((pages + tile_pages - 1) / tile_pages > II_MAX_TILES) tile_pages += 256;
    const uint32_t tiles = pages ? (pages + tile_pages - 1) / tile_pages : 1;
    const size_t words = 1 + (size_t)groups + (size_t)groups * tiles;
    uint32_t* work = ii_buffer(s, words);
    if (!work) return fail(UH_ERR_OOM, "uh_invert_index: work buffer");
    uint32_t* bad = work;
    uint32_t* starts = work + 1;
    uint32_t* tile_count = work + 1 + groups;
    UH_HIP(hipMemsetAsync(bad, 0, 4, s));
    if (!pages) {
        UH_HIP(hipMemsetAsync(totals, 0, (size_t)groups * 4, s));
        UH_HIP(hipStreamSynchronize(s));
        return UH_OK;
    }
    const size_t lds = (size_t)groups * 4;
    UH_LAUNCH(k_ii_count, dim3(tiles), dim3(256), lds, s, dump, pages, max_out, total_pages, pages_per_subtask, groups,
              tile_pages, tiles, tile_count, bad);
    UH_LAUNCH_CHECK("k_ii_count");
    UH_LAUNCH(k_ii_scan, dim3(groups), dim3(256), 0, s, tile_count, tiles, totals);
    UH_LAUNCH_CHECK("k_ii_scan");
    UH_LAUNCH(k_ii_starts, dim3(1), dim3(256), 0, s, (const uint32_t*)totals, groups, starts);
    UH_LAUNCH_CHECK("k_ii_starts");
    UH_LAUNCH(k_ii_scatter, dim3(tiles), dim3(64), lds, s, dump, pages, max_out, total_pages, pages_per_subtask, groups,
              first_page, tile_pages, tiles, (const uint32_t*)tile_count, (const uint32_t*)totals,
              (const uint32_t*)starts, (IiEntry*)out_entries);
    UH_LAUNCH_CHECK("k_ii_scatter");
    g_plain_launches += 4;
    uint32_t nbad = 0;
    UH_HIP(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
    UH_HIP(hipStreamSynchronize(s));
    if (nbad) return fail(UH_ERR_ARG, "uh_invert_index: %u malformed records or links skipped", nbad);
    return UH_OK;
}

extern "C" int uh_prii_rank(const void* entries, uint32_t n, const float* in_ranks, float initial, float damping,
                            uint32_t first_page, float* out, uint32_t out_pages, uint32_t total_pages, void* stream) {
    if (!out_pages) return UH_OK;
    if (!out || (n && !entries)) return fail(UH_ERR_ARG, "uh_prii_rank: null buffer");
    if ((uintptr_t)entries & 3u) return fail(UH_ERR_ARG, "uh_prii_rank: entries must be 4-byte aligned");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    const uint32_t cap = (uint32_t)cu_count() * 8;
    uint32_t zb = (out_pages + 255) / 256;
    if (zb > cap) zb = cap;
    UH_LAUNCH(k_prii_zero, dim3(zb), dim3(256), 0, s, out, out_pages);
    UH_LAUNCH_CHECK("k_prii_zero");
    g_plain_launches++;
    if (!n) return UH_OK;
    uint32_t rb = (n + 255) / 256;
    if (rb > cap) rb = cap;
    UH_LAUNCH(k_prii_rank, dim3(rb), dim3(256), 0, s, (const IiEntry*)entries, n, in_ranks, initial, damping, first_page,
              out, out_pages, total_pages);
    UH_LAUNCH_CHECK("k_prii_rank");
    g_plain_launches++;
    return UH_OK;
}
