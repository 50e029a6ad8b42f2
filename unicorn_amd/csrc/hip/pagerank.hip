// pageRank kernels for gfx950 (replace pageRank_cuda / zeroInit,
// testSuite/pageRank/source/code/src/pageRankCUDA.cu:46-164).
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

int pagerank_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_pagerank_persistent, dim3(cu_count() * 8), dim3(PR_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_pagerank_persistent");
    g_persistent_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_pagerank_pull(const uh_pagerank_args* a, void* stream) {
    if (!a || !a->row_ptr || !a->w || !a->y || !a->w_next || !a->deg)
        return fail(UH_ERR_ARG, "uh_pagerank_pull: null pointer");
    if (a->rows == 0) return UH_OK;
    if (!a->col) return fail(UH_ERR_ARG, "uh_pagerank_pull: null col");
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
