// Dense GEMM for gfx950: row-major C = alpha * A(MxK) * B(KxN) + beta * C.
// Replaces cublasSgemm in matrixMultiplyBlasCUDA.cu:54-77 and luDecompositionCUDA.cu:190.
//
// precision 0 (split-BF16, 3 products): every fp32 operand x is split into hi = RN_bf16(x) and
// lo = RN_bf16(x - hi); C accumulates hi*hi + hi*lo + lo*hi with v_mfma_f32_32x32x16_bf16 in
// fp32. The dropped lo*lo term and the bf16 representation error are below 2^-16 |a||b| per
// product. The split happens while a tile moves from global memory to LDS, so HBM traffic stays
// that of an fp32 GEMM.
// precision 1: native fp32 MFMA (v_mfma_f32_32x32x2_f32).
// precision 2: fp32 VALU kernel, a plain reference used by tests.
//
// Tiling: 256 threads = 2x2 waves; each wave owns TM x TN blocks of 32x32 outputs
// (tile BM = 64*TM by BN = 64*TN), K advances 32 per step through double-buffered LDS.
// Operands are stored K-contiguous in LDS (B is transposed on the way in) so each lane fetches
// its 8 bf16 MFMA operand values with one 16-byte read.
#include "uh_internal.h"

namespace uh {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;

static constexpr int G_THREADS = 256;
static constexpr int G_BK = 32;
static constexpr int G_LDK = G_BK + 8;        // LDS row stride in bf16 (80 bytes, 16B aligned)

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// hi/lo split of two fp32 values into packed bf16 pairs (x in the low half): hi = RN_bf16(v),
// lo = RN_bf16(v - hi), with v_cvt_pk_bf16_f32 (round to nearest even) doing two roundings per
// instruction.
__device__ __forceinline__ void split2_bf16(float x, float y, uint32_t& hi, uint32_t& lo) {
    const f32x2_t v = {x, y};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
    const f32x2_t r = {x - __uint_as_float(hi << 16), y - __uint_as_float(hi & 0xffff0000u)};
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2_t));
}

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// 16-byte global load (p 16-byte aligned)
__device__ __forceinline__ float4 ld_f4(const UH_GLOBAL float* p) {
    const f32x4_t t = *(const UH_GLOBAL f32x4_t*)p;
    return make_float4(t.x, t.y, t.z, t.w);
}

template <int TM, int TN>
struct GemmTile {
    static constexpr int BM = 64 * TM;
    static constexpr int BN = 64 * TN;
    // per stage: A hi, A lo (BM x LDK), B hi, B lo (BN x LDK), in bf16
    static constexpr int STAGE = 2 * BM * G_LDK + 2 * BN * G_LDK;
    static constexpr int LDS_SHORTS = 2 * STAGE;
    static constexpr int A_LOADS = BM * G_BK / 4 / G_THREADS;   // float4 per thread
    static constexpr int B_CHUNKS = BN * G_BK / 8 / G_THREADS;  // 2x4 chunks per thread
};

template <int TM, int TN>
__device__ __forceinline__ void gemm_tile_split3(const uh_gemm_args& g, uint32_t tile, short* lds) {
    using T = GemmTile<TM, TN>;
    const int tiles_n = (g.N + T::BN - 1) / T::BN;
    const int m0 = (int)(tile / (uint32_t)tiles_n) * T::BM;
    const int n0 = (int)(tile % (uint32_t)tiles_n) * T::BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    float4 ra[T::A_LOADS];
    float4 rb[T::B_CHUNKS][2];

    auto load_global = [&](int k0) {
#pragma unroll
        for (int l = 0; l < T::A_LOADS; ++l) {
            const int idx = tid + l * G_THREADS;           // float4 index in the BM x 32 tile
            const int row = idx / (G_BK / 4), kq = (idx % (G_BK / 4)) * 4;
            const int gr = m0 + row, gk = k0 + kq;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gr < g.M) {
                const UH_GLOBAL float* p = as_global(g.A) + (uint64_t)gr * g.lda + gk;
                if (gk + 3 < g.K && ((((uintptr_t)p) & 15u) == 0)) v = ld_f4(p);
                else {
                    if (gk + 0 < g.K) v.x = p[0];
                    if (gk + 1 < g.K) v.y = p[1];
                    if (gk + 2 < g.K) v.z = p[2];
                    if (gk + 3 < g.K) v.w = p[3];
                }
            }
            ra[l] = v;
        }
#pragma unroll
        for (int l = 0; l < T::B_CHUNKS; ++l) {
            const int idx = tid + l * G_THREADS;           // chunk: 2 k-rows x 4 n-cols
            const int kp = (idx / (T::BN / 4)) * 2, nq = (idx % (T::BN / 4)) * 4;
            const int gn = n0 + nq;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int gk = k0 + kp + h;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (gk < g.K) {
                    const UH_GLOBAL float* p = as_global(g.B) + (uint64_t)gk * g.ldb + gn;
                    if (gn + 3 < g.N && ((((uintptr_t)p) & 15u) == 0)) v = ld_f4(p);
                    else {
                        if (gn + 0 < g.N) v.x = p[0];
                        if (gn + 1 < g.N) v.y = p[1];
                        if (gn + 2 < g.N) v.z = p[2];
                        if (gn + 3 < g.N) v.w = p[3];
                    }
                }
                rb[l][h] = v;
            }
        }
    };

    auto store_lds = [&](short* st) {
        short* a_hi = st;
        short* a_lo = st + T::BM * G_LDK;
        short* b_hi = st + 2 * T::BM * G_LDK;
        short* b_lo = b_hi + T::BN * G_LDK;
#pragma unroll
        for (int l = 0; l < T::A_LOADS; ++l) {
            const int idx = tid + l * G_THREADS;
            const int row = idx / (G_BK / 4), kq = (idx % (G_BK / 4)) * 4;
            uint2 hv, lv;
            split2_bf16(ra[l].x, ra[l].y, hv.x, lv.x);
            split2_bf16(ra[l].z, ra[l].w, hv.y, lv.y);
            *(uint2*)(a_hi + row * G_LDK + kq) = hv;
            *(uint2*)(a_lo + row * G_LDK + kq) = lv;
        }
#pragma unroll
        for (int l = 0; l < T::B_CHUNKS; ++l) {
            const int idx = tid + l * G_THREADS;
            const int kp = (idx / (T::BN / 4)) * 2, nq = (idx % (T::BN / 4)) * 4;
            const float v0[4] = {rb[l][0].x, rb[l][0].y, rb[l][0].z, rb[l][0].w};
            const float v1[4] = {rb[l][1].x, rb[l][1].y, rb[l][1].z, rb[l][1].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t hw, lw;
                split2_bf16(v0[e], v1[e], hw, lw);     // k in the low half, k + 1 in the high half
                *(uint32_t*)(b_hi + (nq + e) * G_LDK + kp) = hw;
                *(uint32_t*)(b_lo + (nq + e) * G_LDK + kp) = lw;
            }
        }
    };

    const int ksteps = (g.K + G_BK - 1) / G_BK;
    load_global(0);
    store_lds(lds);
    __syncthreads();
    for (int ks = 0; ks < ksteps; ++ks) {
        short* cur = lds + (ks & 1) * T::STAGE;
        if (ks + 1 < ksteps) load_global((ks + 1) * G_BK);
        const short* a_hi = cur;
        const short* a_lo = cur + T::BM * G_LDK;
        const short* b_hi = cur + 2 * T::BM * G_LDK;
        const short* b_lo = b_hi + T::BN * G_LDK;
#pragma unroll
        for (int kk = 0; kk < G_BK; kk += 16) {
            const int ko = kk + (lane >> 5) * 8;
            bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int row = wm + i * 32 + (lane & 31);
                ah[i] = *(const bf16x8*)(a_hi + row * G_LDK + ko);
                al[i] = *(const bf16x8*)(a_lo + row * G_LDK + ko);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int col = wn + j * 32 + (lane & 31);
                bh[j] = *(const bf16x8*)(b_hi + col * G_LDK + ko);
                bl[j] = *(const bf16x8*)(b_lo + col * G_LDK + ko);
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        }
        if (ks + 1 < ksteps) {
            store_lds(lds + ((ks + 1) & 1) * T::STAGE);
        }
        __syncthreads();
    }

    // epilogue: D[row][col], col = lane % 32, row = (r / 4) * 8 + (lane / 32) * 4 + r % 4
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                const int col = n0 + wn + j * 32 + (lane & 31);
                if (row < g.M && col < g.N) {
                    UH_GLOBAL float* c = as_global(g.C) + (uint64_t)row * g.ldc + col;
                    const float v = g.alpha * acc[i][j][r];
                    *c = g.beta == 0.f ? v : v + g.beta * *c;
                }
            }
}

// fp32 VALU reference: 16x16 threads, each one output element per 64x64 tile pass.
__device__ __forceinline__ void gemm_tile_valu(const uh_gemm_args& g, uint32_t tile) {
    const int tiles_n = (g.N + 63) / 64;
    const int m0 = (int)(tile / (uint32_t)tiles_n) * 64;
    const int n0 = (int)(tile % (uint32_t)tiles_n) * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += G_THREADS) {
        const int row = m0 + e / 64, col = n0 + e % 64;
        if (row >= g.M || col >= g.N) continue;
        float s = 0.f;
        const UH_GLOBAL float* a = as_global(g.A) + (uint64_t)row * g.lda;
        const UH_GLOBAL float* b = as_global(g.B);
        for (int k = 0; k < g.K; ++k) s = fmaf(a[k], b[(uint64_t)k * g.ldb + col], s);
        UH_GLOBAL float* c = as_global(g.C) + (uint64_t)row * g.ldc + col;
        const float v = g.alpha * s;
        *c = g.beta == 0.f ? v : v + g.beta * *c;
    }
}

// Tile configurations selectable per shape class (the autotuner's search space).
struct TileCfg { int tm, tn; };
static constexpr TileCfg kTiles[] = {{2, 2}, {1, 1}, {2, 1}, {1, 2}};
static constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
static constexpr int kMaxLdsShorts = GemmTile<2, 2>::LDS_SHORTS;

__device__ __forceinline__ void gemm_dispatch(const uh_gemm_args& g, uint32_t tile, int cfg, short* lds) {
    if (g.precision == 2) { gemm_tile_valu(g, tile); return; }
    switch (cfg) {
        case 0: gemm_tile_split3<2, 2>(g, tile, lds); break;
        case 1: gemm_tile_split3<1, 1>(g, tile, lds); break;
        case 2: gemm_tile_split3<2, 1>(g, tile, lds); break;
        default: gemm_tile_split3<1, 2>(g, tile, lds); break;
    }
}

__host__ __device__ inline uint32_t gemm_tiles_for(const uh_gemm_args& g, int cfg) {
    int bm = 64, bn = 64;
    if (g.precision != 2) { bm = 64 * kTiles[cfg].tm; bn = 64 * kTiles[cfg].tn; }
    return (uint32_t)(((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn));
}

// The tile configuration travels in the item (stored in the upper bits of precision).
__global__ void __launch_bounds__(G_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm(uh_gemm_args g, int cfg) {
    __shared__ short lds[kMaxLdsShorts];
    // XCD-aware order: blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
    // range of tiles so that tiles sharing A rows share that XCD's L2.
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    uint32_t tile = b;
    const uint32_t per = nb / 8;
    if (per > 0 && b < per * 8) tile = (b % 8) * per + b / 8;
    gemm_dispatch(g, tile, cfg, lds);
}

__global__ void __launch_bounds__(G_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm_persistent(const uint8_t* queue) {
    // 80 KB: two workgroups per CU. The claimed tile index is passed through the first dword of
    // the tile buffers, which the tile itself overwrites only after the second barrier.
    __shared__ short lds[kMaxLdsShorts];
    uint32_t* s_tile = (uint32_t*)lds;
    QueueHeader* h = (QueueHeader*)queue;
    const uh_gemm_args* items = (const uh_gemm_args*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    for (;;) {
        if (threadIdx.x == 0) *s_tile = atomicAdd(&h->next_tile, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(*s_tile);
        __syncthreads();
        if (t >= total) break;
        uint32_t item, local;
        queue_locate(prefix, n_items, t, item, local);
        uh_gemm_args g = uniform_load(items + item);
        const int cfg = g.precision >> 8;
        g.precision &= 0xff;
        gemm_dispatch(g, local, cfg, lds);
        __syncthreads();
    }
}

// ---- tile selection (filled by the autotuner; default by size) ----
struct TuneKey { int m, n, k, p; };
static std::vector<std::pair<TuneKey, int>> g_tuned;
static std::mutex g_tune_mtx;

static int size_class(int v) {   // log2 bucket
    int c = 0;
    while ((1 << (c + 1)) <= v && c < 30) ++c;
    return c;
}

static int pick_tile(const uh_gemm_args& g) {
    {
        std::lock_guard<std::mutex> l(g_tune_mtx);
        for (auto& e : g_tuned)
            if (e.first.m == size_class(g.M) && e.first.n == size_class(g.N) && e.first.k == size_class(g.K) &&
                e.first.p == g.precision)
                return e.second;
    }
    // default: big tiles when they fill the chip, smaller ones otherwise
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    return t128 >= cu_count() ? 0 : 1;
}

uint32_t gemm_tiles(const uh_gemm_args& g) { return gemm_tiles_for(g, g.precision >> 8); }

int gemm_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_gemm_persistent, dim3(cu_count() * 2), dim3(G_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_gemm_persistent");
    g_persistent_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_gemm(const uh_gemm_args* args, void* stream) {
    if (!args || !args->A || !args->B || !args->C) return fail(UH_ERR_ARG, "uh_gemm: null pointer");
    const uh_gemm_args& g = *args;
    if (g.M < 0 || g.N < 0 || g.K < 0) return fail(UH_ERR_ARG, "uh_gemm: negative size");
    if (g.precision < 0 || g.precision > 2) return fail(UH_ERR_ARG, "uh_gemm: precision %d", g.precision);
    if (g.precision == 1) return fail(UH_ERR_ARG, "uh_gemm: fp32 MFMA path not built");
    if (g.M == 0 || g.N == 0) return UH_OK;
    if ((uint64_t)g.lda < (uint64_t)g.K || (uint64_t)g.ldb < (uint64_t)g.N || (uint64_t)g.ldc < (uint64_t)g.N)
        return fail(UH_ERR_ARG, "uh_gemm: leading dimension too small");
    const int cfg = pick_tile(g);
    const uint32_t tiles = gemm_tiles_for(g, cfg);
    if (Batch* b = as_batch(stream)) {
        uh_gemm_args item = g;
        item.precision = g.precision | (cfg << 8);
        record_item(b, FAM_GEMM, &item, sizeof(item), tiles);
        return UH_OK;
    }
    UH_LAUNCH(k_gemm, dim3(tiles), dim3(G_THREADS), 0, (hipStream_t)stream, g, cfg);
    UH_LAUNCH_CHECK("k_gemm");
    g_plain_launches++;
    return UH_OK;
}

extern "C" int uh_gemm_set_tile(int M, int N, int K, int precision, int tile_id) {
    if (tile_id < 0 || tile_id >= kNumTiles) return fail(UH_ERR_ARG, "uh_gemm_set_tile: tile %d", tile_id);
    std::lock_guard<std::mutex> l(g_tune_mtx);
    TuneKey k{size_class(M), size_class(N), size_class(K), precision};
    for (auto& e : g_tuned)
        if (e.first.m == k.m && e.first.n == k.n && e.first.k == k.k && e.first.p == k.p) { e.second = tile_id; return UH_OK; }
    g_tuned.push_back({k, tile_id});
    return UH_OK;
}

extern "C" int uh_gemm_tile_count(int* n) { *n = kNumTiles; return UH_OK; }

extern "C" int uh_gemm_autotune(int M, int N, int K, int precision, int reps, int force, int* tile_id, float* ms) {
    if (M <= 0 || N <= 0 || K <= 0 || precision != 0 || reps < 1) return fail(UH_ERR_ARG, "uh_gemm_autotune: arguments");
    const TuneKey key{size_class(M), size_class(N), size_class(K), precision};
    if (!force) {
        std::lock_guard<std::mutex> l(g_tune_mtx);
        for (auto& e : g_tuned)
            if (e.first.m == key.m && e.first.n == key.n && e.first.k == key.k && e.first.p == key.p) {
                if (tile_id) *tile_id = e.second;
                if (ms) *ms = -1.f;
                return UH_OK;
            }
    }
    float* buf = nullptr;
    const size_t elems = (size_t)M * K + (size_t)K * N + (size_t)M * N;
    UH_HIP(hipMalloc((void**)&buf, elems * sizeof(float)));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = -1, rc = UH_OK;
    float best_ms = 0.f;
    auto run = [&]() -> int {
        UH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        UH_HIP(hipEventCreate(&e0));
        UH_HIP(hipEventCreate(&e1));
        UH_HIP(hipMemsetAsync(buf, 0, elems * sizeof(float), s));
        uh_gemm_args g{};
        g.A = buf; g.B = buf + (size_t)M * K; g.C = buf + (size_t)M * K + (size_t)K * N;
        g.lda = (uint64_t)K; g.ldb = (uint64_t)N; g.ldc = (uint64_t)N;
        g.M = M; g.N = N; g.K = K; g.precision = precision; g.alpha = 1.f; g.beta = 0.f;
        for (int cfg = 0; cfg < kNumTiles; ++cfg) {
            const uint32_t tiles = gemm_tiles_for(g, cfg);
            UH_LAUNCH(k_gemm, dim3(tiles), dim3(G_THREADS), 0, s, g, cfg);
            UH_LAUNCH_CHECK("k_gemm (autotune)");
            UH_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) UH_LAUNCH(k_gemm, dim3(tiles), dim3(G_THREADS), 0, s, g, cfg);
            UH_LAUNCH_CHECK("k_gemm (autotune)");
            UH_HIP(hipEventRecord(e1, s));
            UH_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            UH_HIP(hipEventElapsedTime(&t, e0, e1));
            t /= (float)reps;
            if (best < 0 || t < best_ms) { best = cfg; best_ms = t; }
        }
        return UH_OK;
    };
    rc = run();
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    hipFree(buf);
    if (rc) return rc;
    uh_gemm_set_tile(M, N, K, precision, best);
    if (tile_id) *tile_id = best;
    if (ms) *ms = best_ms;
    return UH_OK;
}
