// pageRank over the binned (propagation-blocked) graph layout, for gfx950 (replaces
// pageRank_cuda / zeroInit, testSuite/pageRank/source/code/src/pageRankCUDA.cu:46-164).
//
// Why binned. A pull over an in-link list gathers w[src] for random sources: one L2 request per
// link, and ~100 M links per iteration at 10 M pages make that request rate the bound (the
// source-blocked pull in pagerank.hip, which keeps an eighth of w per XCD, is limited by exactly
// this). Here every random access goes to LDS and HBM is only streamed:
//   - pages form units of at most UH_PRB_UNIT_MAX consecutive pages; a unit is both a
//     destination *bin* and a source *chunk*;
//   - the links' contributions ("values") live in a bin-major array: bin b holds the values of
//     every link into b, grouped by source chunk (a *run* (c, b) is the links c -> b);
//   - a unit u: (phase 2) streams the values of bin u with their destinations (dst16) and adds
//     them into an LDS accumulator of the unit's rows, writes y and turns it into the unit's w in
//     LDS; then (phase 1) for every run (u, b) streams the sources (src16) of chunk u and writes
//     w[src] into the values of the next iteration, at the run's place in bin b.
// The iteration's values are read once and the next iteration's written once, both contiguous.
//
// Determinism. The accumulator is fixed point: each value v >= 0 adds floor(v * 2^s) as an
// unsigned 64-bit integer (ds_add_u64), so the sum does not depend on the order of the atomic
// adds; y = (float)((double)acc * 2^-s). s is chosen by the caller so that no row sum can reach
// 2^62 (every row sum is at most N * x0 when damping <= 1). The CPU callback does the same integer
// arithmetic, so both produce the same bits.
#include "uh_internal.h"

namespace uh {

static constexpr int PRB_THREADS = 1024;
static constexpr int PRB_WAVES = PRB_THREADS / 64;
static constexpr uint32_t PRB_UNIT = UH_PRB_UNIT_MAX;
static_assert(PRB_UNIT % PRB_THREADS == 0, "rows per thread");
static constexpr int PRB_ROWS_PER_THREAD = PRB_UNIT / PRB_THREADS;
typedef float prb_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t prb_u2 __attribute__((ext_vector_type(2)));

// floor(v * 2^s) for a finite v >= 0 whose scaled value is below 2^63 (exact: the float's
// integer significand shifted).
__device__ __forceinline__ uint64_t prb_fixed(float v, int s) {
    const uint32_t b = __float_as_uint(v);
    int e = (int)(b >> 23);
    uint64_t m = b & 0x7fffffu;
    if (e) m |= 0x800000u; else e = 1;
    const int sh = e - 150 + s;                   // v = m * 2^(e - 150)
    return sh >= 0 ? (m << sh) : (sh > -64 ? (m >> -sh) : 0ull);
}

// One unit of an item. acc is the workgroup's LDS (PRB_UNIT 64-bit words), reused as floats for w.
__device__ __forceinline__ void prb_unit(const uh_pagerank_binned_args& a, uint32_t u, uint64_t* acc) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const UH_GLOBAL uint32_t* unit_start = as_global(a.unit_start);
    const uint32_t p0 = __builtin_amdgcn_readfirstlane(unit_start[u]);
    const uint32_t rows = __builtin_amdgcn_readfirstlane(unit_start[u + 1]) - p0;
    float* w = (float*)acc;
    float wr[PRB_ROWS_PER_THREAD];
    if (a.kind == 0) {
        // first iteration: every page has rank x0
#pragma unroll
        for (int j = 0; j < PRB_ROWS_PER_THREAD; ++j) {
            const uint32_t r = tid + j * PRB_THREADS;
            wr[j] = 0.f;
            if (r < rows) {
                as_global(a.y)[p0 + r] = a.x0;
                const uint32_t d = as_global(a.deg)[p0 + r];
                wr[j] = d ? (a.damping * a.x0) / (float)d : 0.f;
            }
        }
    } else {
        // phase 2: the bin's values into the fixed-point accumulator
        for (uint32_t r = tid; r < rows; r += PRB_THREADS) acc[r] = 0ull;
        __syncthreads();
        const UH_GLOBAL uint32_t* bin_start = as_global(a.bin_start);
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(bin_start[u]);
        const uint32_t b1 = __builtin_amdgcn_readfirstlane(bin_start[u + 1]);   // both multiples of 4
        const UH_GLOBAL prb_f4* vin = (const UH_GLOBAL prb_f4*)as_global(a.values_in + b0);
        const UH_GLOBAL prb_u2* dst = (const UH_GLOBAL prb_u2*)as_global(a.dst16 + b0);
        const uint32_t n4 = (b1 - b0) >> 2;
        const int s = a.scale_exp;
        for (uint32_t q = tid; q < n4; q += 2 * PRB_THREADS) {
            const bool two = q + PRB_THREADS < n4;
            const prb_f4 v0 = __builtin_nontemporal_load(vin + q);
            const prb_u2 d0 = __builtin_nontemporal_load(dst + q);
            prb_f4 v1 = {0.f, 0.f, 0.f, 0.f};
            prb_u2 d1 = {0u, 0u};
            if (two) {
                v1 = __builtin_nontemporal_load(vin + q + PRB_THREADS);
                d1 = __builtin_nontemporal_load(dst + q + PRB_THREADS);
            }
            atomicAdd((unsigned long long*)&acc[d0.x & 0xffffu], (unsigned long long)prb_fixed(v0.x, s));
            atomicAdd((unsigned long long*)&acc[d0.x >> 16], (unsigned long long)prb_fixed(v0.y, s));
            atomicAdd((unsigned long long*)&acc[d0.y & 0xffffu], (unsigned long long)prb_fixed(v0.z, s));
            atomicAdd((unsigned long long*)&acc[d0.y >> 16], (unsigned long long)prb_fixed(v0.w, s));
            if (two) {
                atomicAdd((unsigned long long*)&acc[d1.x & 0xffffu], (unsigned long long)prb_fixed(v1.x, s));
                atomicAdd((unsigned long long*)&acc[d1.x >> 16], (unsigned long long)prb_fixed(v1.y, s));
                atomicAdd((unsigned long long*)&acc[d1.y & 0xffffu], (unsigned long long)prb_fixed(v1.z, s));
                atomicAdd((unsigned long long*)&acc[d1.y >> 16], (unsigned long long)prb_fixed(v1.w, s));
            }
        }
        __syncthreads();
        const double unscale = __builtin_ldexp(1.0, -s);
#pragma unroll
        for (int j = 0; j < PRB_ROWS_PER_THREAD; ++j) {
            const uint32_t r = tid + j * PRB_THREADS;
            wr[j] = 0.f;
            if (r < rows) {
                const float y = (float)((double)(long long)acc[r] * unscale);
                as_global(a.y)[p0 + r] = y;
                const uint32_t d = as_global(a.deg)[p0 + r];
                wr[j] = d ? (a.damping * y) / (float)d : 0.f;
            }
        }
    }
    if (!a.values_out) {
        __syncthreads();            // acc is reused by the next unit
        return;
    }
    __syncthreads();                // every row's accumulator was read
#pragma unroll
    for (int j = 0; j < PRB_ROWS_PER_THREAD; ++j) {
        const uint32_t r = tid + j * PRB_THREADS;
        if (r < rows) w[r] = wr[j];
    }
    __syncthreads();
    // phase 1: chunk u's runs, one wave per run; each run's sources and value slots are contiguous
    const UH_GLOBAL uint32_t* run_ptr = as_global(a.run_ptr);
    const UH_GLOBAL uint32_t* run_off = as_global(a.run_off);
    const UH_GLOBAL uint32_t* run_pos = as_global(a.run_pos);
    const UH_GLOBAL uint16_t* src = as_global(a.src16);
    UH_GLOBAL float* vout = as_global(a.values_out);
    const uint32_t r0 = __builtin_amdgcn_readfirstlane(run_ptr[u]);
    const uint32_t r1 = __builtin_amdgcn_readfirstlane(run_ptr[u + 1]);
    for (uint32_t r = r0 + wave; r < r1; r += PRB_WAVES) {
        const uint32_t o0 = __builtin_amdgcn_readfirstlane(run_off[r]);
        const uint32_t o1 = __builtin_amdgcn_readfirstlane(run_off[r + 1]);
        const uint32_t pos = __builtin_amdgcn_readfirstlane(run_pos[r]);
        for (uint32_t k = o0 + lane; k < o1; k += 256) {
            uint32_t sv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t kk = k + 64 * j;
                sv[j] = kk < o1 ? (uint32_t)__builtin_nontemporal_load(src + kk) : 0u;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t kk = k + 64 * j;
                if (kk < o1) __builtin_nontemporal_store(w[sv[j]], vout + pos + (kk - o0));
            }
        }
    }
    __syncthreads();                // w is overwritten by the next unit's accumulator
}

__global__ void __launch_bounds__(PRB_THREADS) k_pagerank_binned(uh_pagerank_binned_args a) {
    extern __shared__ uint64_t prb_acc[];
    prb_unit(a, a.unit0 + blockIdx.x, prb_acc);
}

// Batched: tiles are the units of the items, claimed from the queue's counter.
__global__ void __launch_bounds__(PRB_THREADS) k_pagerank_binned_persistent(const uint8_t* queue) {
    extern __shared__ uint64_t prb_acc[];
    __shared__ uint32_t s_t;
    QueueHeader* h = (QueueHeader*)queue;
    const uh_pagerank_binned_args* items = (const uh_pagerank_binned_args*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(&h->next_tile, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        if (t >= total) break;
        uint32_t item, local;
        queue_locate(prefix, n_items, t, item, local);
        const uh_pagerank_binned_args a = uniform_load(items + item);
        prb_unit(a, a.unit0 + local, prb_acc);
    }
}

static constexpr size_t kPrbLds = (size_t)PRB_UNIT * sizeof(uint64_t);

static int prb_lds_ok() {
    static int ok = -1;
    if (ok < 0) {
        ok = hipFuncSetAttribute((const void*)k_pagerank_binned, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kPrbLds) == hipSuccess &&
             hipFuncSetAttribute((const void*)k_pagerank_binned_persistent,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPrbLds) == hipSuccess ? 1 : 0;
    }
    return ok;
}

int pagerank_binned_persistent(const uint8_t* queue, hipStream_t s) {
    if (!prb_lds_ok()) return fail(UH_ERR_HIP, "k_pagerank_binned_persistent: %zu bytes of LDS refused", kPrbLds);
    UH_LAUNCH(k_pagerank_binned_persistent, dim3(cu_count()), dim3(PRB_THREADS), kPrbLds, s, queue);
    UH_LAUNCH_CHECK("k_pagerank_binned_persistent");
    g_persistent_launches++;
    return UH_OK;
}

int pagerank_binned_prepare() { return prb_lds_ok() ? UH_OK : fail(UH_ERR_HIP, "binned pageRank: LDS size refused"); }

}  // namespace uh

using namespace uh;

extern "C" int uh_pagerank_binned(const uh_pagerank_binned_args* a, void* stream) {
    if (!a || !a->unit_start || !a->deg || !a->y)
        return fail(UH_ERR_ARG, "uh_pagerank_binned: null pointer");
    if (a->kind != 0 && a->kind != 1) return fail(UH_ERR_ARG, "uh_pagerank_binned: kind must be 0 or 1");
    // values_in / dst16 (run_pos / src16) may be NULL only when the units' bins (runs) are empty:
    // the kernel reads them per slot (run) of the units
    if (a->kind == 1 && !a->bin_start) return fail(UH_ERR_ARG, "uh_pagerank_binned: iterate kind without bin_start");
    if (a->values_out && (!a->run_ptr || !a->run_off))
        return fail(UH_ERR_ARG, "uh_pagerank_binned: values_out without run_ptr / run_off");
    if (a->kind == 1 && (((uintptr_t)a->values_in & 15u) || ((uintptr_t)a->dst16 & 7u)))
        return fail(UH_ERR_ARG, "uh_pagerank_binned: values_in (16 bytes) and dst16 (8 bytes) misaligned");
    if (a->scale_exp < 0 || a->scale_exp > 100) return fail(UH_ERR_ARG, "uh_pagerank_binned: scale exponent out of range");
    if (a->units == 0) return UH_OK;
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_PRBIN, a, sizeof(*a), a->units);
        return UH_OK;
    }
    if (int rc = pagerank_binned_prepare()) return rc;
    UH_LAUNCH(k_pagerank_binned, dim3(a->units), dim3(PRB_THREADS), kPrbLds, (hipStream_t)stream, *a);
    UH_LAUNCH_CHECK("k_pagerank_binned");
    g_plain_launches++;
    return UH_OK;
}
