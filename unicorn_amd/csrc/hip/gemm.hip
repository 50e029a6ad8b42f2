// Dense GEMM for gfx950: row-major C = alpha * A(MxK) * B(KxN) + beta * C.
// Replaces cublasSgemm in matrixMultiplyBlasCUDA.cu:54-77 and luDecompositionCUDA.cu:190.
//
// precision 0 (split-BF16, 3 products): every fp32 operand x is split into hi = RN_bf16(x) and
// lo = RN_bf16(x - hi); C accumulates lo*hi + hi*lo + hi*hi with v_mfma_f32_32x32x16_bf16 in
// fp32. The dropped lo*lo term and the bf16 representation error are below 2^-16 |a||b| per
// product. The split happens while a tile moves from global memory to LDS (v_cvt_pk_bf16_f32),
// so HBM traffic stays that of an fp32 GEMM.
// precision 1: native fp32 MFMA (not built; rejected by uh_gemm).
// precision 2: fp32 VALU kernel, a plain reference used by tests.
//
// Tiling: 256 threads = 2x2 waves; each wave owns TM x TN blocks of 32x32 outputs
// (tile BM = 64*TM by BN = 64*TN); K advances BK per stage through double-buffered LDS, and the
// next stage's global loads are in flight during the current stage's MFMAs. Operands are stored
// K-contiguous in LDS (B is transposed on the way in) so each lane fetches its 8 bf16 MFMA operand
// values with one 16-byte read. Each tile configuration is its own kernel, with LDS sized for it.
#include "uh_internal.h"

namespace uh {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

static constexpr int G_THREADS = 256;

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// hi/lo split of two fp32 values into packed bf16 pairs (x in the low half): hi = RN_bf16(v),
// lo = RN_bf16(v - hi), with v_cvt_pk_bf16_f32 (round to nearest even) doing two roundings per
// instruction.
__device__ __forceinline__ void split2_bf16(float x, float y, uint32_t& hi, uint32_t& lo) {
    const f32x2_t v = {x, y};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
    const f32x2_t r = {x - __uint_as_float(hi << 16), y - __uint_as_float(hi & 0xffff0000u)};
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2_t));
}

// 16-byte global load (p 16-byte aligned)
__device__ __forceinline__ float4 ld_f4(const UH_GLOBAL float* p) {
    const f32x4_t t = *(const UH_GLOBAL f32x4_t*)p;
    return make_float4(t.x, t.y, t.z, t.w);
}

// Tile configurations (the autotuner's search space): waves' 32x32 blocks and the K depth of a
// stage. LDS per workgroup: 2 stages x (hi + lo) x (BM + BN) x (BK + 8) x 2 bytes.
struct TileCfg { int tm, tn, bk; };
static constexpr TileCfg kTiles[] = {
    {2, 2, 64},   // 0: 128 x 128, BK 64 (144 KB LDS, 1 workgroup per CU)
    {1, 1, 64},   // 1:  64 x  64, BK 64 (72 KB, 2 per CU)
    {2, 1, 64},   // 2: 128 x  64, BK 64 (108 KB)
    {1, 2, 64},   // 3:  64 x 128, BK 64 (108 KB)
    {2, 2, 32},   // 4: 128 x 128, BK 32 (80 KB, 2 per CU)
    {1, 1, 32},   // 5:  64 x  64, BK 32 (40 KB)
    {2, 1, 32},   // 6: 128 x  64, BK 32 (60 KB, 2 per CU)
    {1, 2, 32},   // 7:  64 x 128, BK 32 (60 KB, 2 per CU)
};
static constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
static constexpr int G_VALU = kNumTiles;          // variant of the precision-2 reference path
static constexpr int G_GENERIC = kNumTiles + 1;   // variant that dispatches per item

template <int TM, int TN, int BK>
struct GemmTile {
    static constexpr int BM = 64 * TM;
    static constexpr int BN = 64 * TN;
    static constexpr int LDK = BK + 8;            // LDS row stride in bf16 (16-byte multiple)
    // per stage: A hi, A lo (BM x LDK), B hi, B lo (BN x LDK), in bf16
    static constexpr int STAGE = 2 * BM * LDK + 2 * BN * LDK;
    static constexpr int LDS_SHORTS = 2 * STAGE;
    static constexpr int A_LOADS = BM * BK / 4 / G_THREADS;   // float4 per thread
    // B: thread t stages column t % BN of the tile, K run (t / BN) * KB .. + KB - 1
    static constexpr int KB = BK * BN / G_THREADS;
    static_assert(BM * BK % (4 * G_THREADS) == 0, "A tile / threads");
    static_assert(G_THREADS % BN == 0 && KB % 8 == 0, "B tile / threads");
};

template <int CFG>
struct Cfg {
    static constexpr int TM = kTiles[CFG].tm, TN = kTiles[CFG].tn, BK = kTiles[CFG].bk;
    using T = GemmTile<TM, TN, BK>;
};

template <int TM, int TN, int BK>
__device__ __forceinline__ void gemm_tile_split3(const uh_gemm_args& g, uint32_t tile, short* lds) {
    using T = GemmTile<TM, TN, BK>;
    constexpr int LDK = T::LDK;
    const int tiles_n = (g.N + T::BN - 1) / T::BN;
    const int m0 = (int)(tile / (uint32_t)tiles_n) * T::BM;
    const int n0 = (int)(tile % (uint32_t)tiles_n) * T::BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32 * TM, wn = (wave & 1) * 32 * TN;
    const UH_GLOBAL float* A = as_global(g.A);
    const UH_GLOBAL float* B = as_global(g.B);

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    float4 ra[T::A_LOADS];
    float rb[T::KB];
    const int bn = tid % T::BN, bk = (tid / T::BN) * T::KB;     // this thread's B column and K run

    // Interior tiles (no edge in M, N or K, 16-byte aligned A rows, 32-bit buffer offsets) load
    // through buffer resources: the per-lane offsets are fixed for the whole tile and the K step
    // is a scalar offset, so the loads cost no vector address arithmetic.
    const bool full = m0 + T::BM <= g.M && n0 + T::BN <= g.N && g.K % BK == 0 &&
                      (((uintptr_t)g.A | (g.lda * 4)) & 15u) == 0 &&
                      (uint64_t)T::BM * g.lda * 4 < (1ull << 31) && (uint64_t)g.K * g.ldb * 4 < (1ull << 31);
    __amdgpu_buffer_rsrc_t rsrc_a = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(g.A + (uint64_t)m0 * g.lda), 0, (int)(full ? T::BM * g.lda * 4 : 0), 0x00020000);
    __amdgpu_buffer_rsrc_t rsrc_b = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(g.B + n0), 0, (int)(full ? g.K * g.ldb * 4 : 0), 0x00020000);
    const int row0 = tid / (BK / 4), kq0 = (tid % (BK / 4)) * 4;   // A float4: row, k of l = 0
    constexpr int A_ROW_STEP = G_THREADS / (BK / 4);                // rows between l and l + 1
    const uint32_t voff_a = (uint32_t)((row0 * g.lda + kq0) * 4);
    const uint32_t voff_b = (uint32_t)((bk * g.ldb + bn) * 4);
    const uint32_t bstep = (uint32_t)(__builtin_amdgcn_readfirstlane((int)g.ldb) * 4);

    auto load_full = [&](int k0) {
#pragma unroll
        for (int l = 0; l < T::A_LOADS; ++l) {
            const f32x4_t v = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(
                rsrc_a, voff_a + (uint32_t)(l * A_ROW_STEP * g.lda * 4), (uint32_t)k0 * 4u, 0));
            ra[l] = make_float4(v.x, v.y, v.z, v.w);
        }
#pragma unroll
        for (int j = 0; j < T::KB; ++j)
            rb[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                rsrc_b, voff_b, (uint32_t)(k0 + j) * bstep, 0));
    };

    auto load_edge = [&](int k0) {
#pragma unroll
        for (int l = 0; l < T::A_LOADS; ++l) {
            const int idx = tid + l * G_THREADS;           // float4 index in the BM x BK tile
            const int row = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
            const int gr = m0 + row, gk = k0 + kq;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gr < g.M) {
                const UH_GLOBAL float* p = A + (uint64_t)gr * g.lda + gk;
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
        // B: one column per lane, consecutive lanes on consecutive columns (coalesced rows)
        const int gn = n0 + bn;
        const UH_GLOBAL float* pb = B + (uint64_t)(k0 + bk) * g.ldb + gn;
#pragma unroll
        for (int j = 0; j < T::KB; ++j) {
            const int gk = k0 + bk + j;
            rb[j] = (gn < g.N && gk < g.K) ? pb[(uint64_t)j * g.ldb] : 0.f;
        }
    };

    auto store_lds = [&](short* st) {
        short* a_hi = st;
        short* a_lo = st + T::BM * LDK;
        short* b_hi = st + 2 * T::BM * LDK;
        short* b_lo = b_hi + T::BN * LDK;
#pragma unroll
        for (int l = 0; l < T::A_LOADS; ++l) {
            const int idx = tid + l * G_THREADS;
            const int row = idx / (BK / 4), kq = (idx % (BK / 4)) * 4;
            uint2 hv, lv;
            split2_bf16(ra[l].x, ra[l].y, hv.x, lv.x);
            split2_bf16(ra[l].z, ra[l].w, hv.y, lv.y);
            *(uint2*)(a_hi + row * LDK + kq) = hv;
            *(uint2*)(a_lo + row * LDK + kq) = lv;
        }
        // B^T row `bn`, K run `bk`: 16-byte writes. Consecutive lanes write consecutive LDS rows,
        // LDK/2 dwords apart (20 or 36), which puts the 8 lanes of each 16-byte write phase on
        // distinct banks
#pragma unroll
        for (int j = 0; j < T::KB; j += 8) {
            uint4 hv, lv;
            split2_bf16(rb[j + 0], rb[j + 1], hv.x, lv.x);     // k in the low half, k + 1 high
            split2_bf16(rb[j + 2], rb[j + 3], hv.y, lv.y);
            split2_bf16(rb[j + 4], rb[j + 5], hv.z, lv.z);
            split2_bf16(rb[j + 6], rb[j + 7], hv.w, lv.w);
            *(uint4*)(b_hi + bn * LDK + bk + j) = hv;
            *(uint4*)(b_lo + bn * LDK + bk + j) = lv;
        }
    };

    // One 16-deep K slice of the stage at `cur`. Each accumulator takes its three products in the
    // order lo*hi, hi*lo, hi*hi; the loops run over the accumulators innermost so that
    // consecutive MFMAs are independent.
    auto mfma_slice = [&](const short* cur, int kk) {
        const short* a_hi = cur;
        const short* a_lo = cur + T::BM * LDK;
        const short* b_hi = cur + 2 * T::BM * LDK;
        const short* b_lo = b_hi + T::BN * LDK;
        const int ko = kk + (lane >> 5) * 8;
        bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int row = wm + i * 32 + (lane & 31);
            ah[i] = *(const bf16x8*)(a_hi + row * LDK + ko);
            al[i] = *(const bf16x8*)(a_lo + row * LDK + ko);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = wn + j * 32 + (lane & 31);
            bh[j] = *(const bf16x8*)(b_hi + col * LDK + ko);
            bl[j] = *(const bf16x8*)(b_lo + col * LDK + ko);
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    };

    // Pipeline: at step ks the registers hold stage ks+1 (loaded during step ks-1); they are split
    // into the free LDS buffer, the loads of stage ks+2 are issued, then stage ks is multiplied.
    // Each load thus has a whole step of MFMAs to arrive.
    auto load_global = [&](int k0) {
        if (full) load_full(k0);
        else load_edge(k0);
    };

    const int ksteps = (g.K + BK - 1) / BK;
    load_global(0);
    store_lds(lds);
    if (ksteps > 1) load_global(BK);
    __syncthreads();
    for (int ks = 0; ks < ksteps; ++ks) {
        const short* cur = lds + (ks & 1) * T::STAGE;
        if (ks + 1 < ksteps) {
            store_lds(lds + ((ks + 1) & 1) * T::STAGE);
            if (ks + 2 < ksteps) load_global((ks + 2) * BK);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) mfma_slice(cur, kk);
        __syncthreads();
    }

    // epilogue: D[row][col], col = lane % 32, row = (r / 4) * 8 + (lane / 32) * 4 + r % 4
    UH_GLOBAL float* C = as_global(g.C);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                const int col = n0 + wn + j * 32 + (lane & 31);
                if (row < g.M && col < g.N) {
                    UH_GLOBAL float* c = C + (uint64_t)row * g.ldc + col;
                    const float v = g.alpha * acc[i][j][r];
                    *c = g.beta == 0.f ? v : v + g.beta * *c;
                }
            }
}

// fp32 VALU reference: each thread computes outputs of a 64x64 tile, one at a time.
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

// The variant an item runs on: its tile configuration (precision 0) or the VALU path.
__host__ __device__ inline int gemm_variant_of(int precision, int cfg) { return precision == 2 ? G_VALU : cfg; }

template <int V> struct VariantLds { static constexpr int SHORTS = Cfg<V>::T::LDS_SHORTS; };
template <> struct VariantLds<G_VALU> { static constexpr int SHORTS = 2; };
template <> struct VariantLds<G_GENERIC> { static constexpr int SHORTS = GemmTile<2, 2, 64>::LDS_SHORTS; };

template <int V>
__device__ __forceinline__ void gemm_run(const uh_gemm_args& g, int cfg, uint32_t tile, short* lds) {
    if constexpr (V == G_VALU) {
        gemm_tile_valu(g, tile);
    } else if constexpr (V == G_GENERIC) {
        if (g.precision == 2) { gemm_tile_valu(g, tile); return; }
        switch (cfg) {
            case 0: gemm_tile_split3<2, 2, 64>(g, tile, lds); break;
            case 1: gemm_tile_split3<1, 1, 64>(g, tile, lds); break;
            case 2: gemm_tile_split3<2, 1, 64>(g, tile, lds); break;
            case 3: gemm_tile_split3<1, 2, 64>(g, tile, lds); break;
            case 4: gemm_tile_split3<2, 2, 32>(g, tile, lds); break;
            case 6: gemm_tile_split3<2, 1, 32>(g, tile, lds); break;
            case 7: gemm_tile_split3<1, 2, 32>(g, tile, lds); break;
            default: gemm_tile_split3<1, 1, 32>(g, tile, lds); break;
        }
    } else {
        gemm_tile_split3<Cfg<V>::TM, Cfg<V>::TN, Cfg<V>::BK>(g, tile, lds);
    }
}

__host__ __device__ inline uint32_t gemm_tiles_for(const uh_gemm_args& g, int cfg) {
    int bm = 64, bn = 64;
    if (g.precision != 2) { bm = 64 * kTiles[cfg].tm; bn = 64 * kTiles[cfg].tn; }
    return (uint32_t)(((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn));
}

// Plain launch of one product (one tile per workgroup).
template <int V>
__global__ void __launch_bounds__(G_THREADS) k_gemm(uh_gemm_args g, int cfg) {
    __shared__ short lds[VariantLds<V>::SHORTS];
    // XCD-aware order: blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
    // range of tiles so that tiles sharing A rows share that XCD's L2.
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    uint32_t tile = b;
    const uint32_t per = nb / 8;
    if (per > 0 && b < per * 8) tile = (b % 8) * per + b / 8;
    gemm_run<V>(g, cfg, tile, lds);
}

// Persistent launch over a work queue of products. The claimed tile index is passed through the
// first dword of the tile buffers, which the tile itself overwrites only after the second barrier.
// The tile configuration travels in the item (upper bits of precision).
template <int V>
__global__ void __launch_bounds__(G_THREADS) k_gemm_persistent(const uint8_t* queue) {
    __shared__ short lds[VariantLds<V>::SHORTS];
    // the claim is passed through the tile buffer itself: configuration 4 uses exactly half
    // of a CU's LDS, and any further shared word would leave room for one workgroup, not two
    uint32_t* s_tile = (uint32_t*)lds;
    QueueHeader* h = (QueueHeader*)queue;
    const QueueHeader hd = uniform_load(h);
    const uh_gemm_args* items = (const uh_gemm_args*)(queue + hd.items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + hd.prefix_offset);
    const uint32_t total = hd.total_tiles, n_items = hd.n_items;
    // The tile range is cut into 8 contiguous parts, part x claimed through counter x (its own
    // cache line of the header). A workgroup claims from its XCD's part first, so an XCD runs
    // whole items and re-reads their A and B panels from its own L2; then it helps the other parts.
    const uint32_t x0 = __builtin_amdgcn_readfirstlane(xcc_id());
    for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t x = (x0 + i) & 7u;
        const uint32_t lo = (uint32_t)(((uint64_t)total * x) / 8), hi = (uint32_t)(((uint64_t)total * (x + 1)) / 8);
        for (;;) {
            if (threadIdx.x == 0) *s_tile = atomicAdd(&h->next_blk[x * 32], 1u);
            __syncthreads();
            const uint32_t k = __builtin_amdgcn_readfirstlane(*s_tile);
            __syncthreads();
            if (k >= hi - lo) break;
            const uint32_t t = lo + k;
            uint32_t item, local;
            queue_locate(prefix, n_items, t, item, local);
            uh_gemm_args g = uniform_load(items + item);
            const int cfg = g.precision >> 8;
            g.precision &= 0xff;
            gemm_run<V>(g, cfg, local, lds);
            __syncthreads();
        }
    }
}

typedef void (*GemmPlainFn)(uh_gemm_args, int);
typedef void (*GemmPersistentFn)(const uint8_t*);
static const GemmPlainFn kGemmPlain[kNumTiles + 1] = {
    k_gemm<0>, k_gemm<1>, k_gemm<2>, k_gemm<3>, k_gemm<4>, k_gemm<5>, k_gemm<6>, k_gemm<7>, k_gemm<G_VALU>};
static const GemmPersistentFn kGemmPersistent[kNumTiles + 2] = {
    k_gemm_persistent<0>, k_gemm_persistent<1>, k_gemm_persistent<2>, k_gemm_persistent<3>,
    k_gemm_persistent<4>, k_gemm_persistent<5>, k_gemm_persistent<6>, k_gemm_persistent<7>,
    k_gemm_persistent<G_VALU>, k_gemm_persistent<G_GENERIC>};
static_assert(kNumTiles == 8, "kernel tables list every tile configuration");

// Workgroups per CU of each persistent variant (occupancy queries, cached).
static int g_gemm_wg[kNumTiles + 2] = {};

int gemm_wg_per_cu() {
    if (!g_gemm_wg[0]) {
        for (int v = 0; v < kNumTiles + 2; ++v) {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kGemmPersistent[v], G_THREADS, 0) != hipSuccess || b < 1)
                b = 1;
            g_gemm_wg[v] = b;
        }
    }
    return g_gemm_wg[0];
}

int gemm_variant_for(uint32_t kinds) {
    return (kinds && (kinds & (kinds - 1)) == 0) ? __builtin_ctz(kinds) : G_GENERIC;
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

int gemm_persistent(const uint8_t* queue, hipStream_t s, uint32_t kinds) {
    const int v = gemm_variant_for(kinds);
    gemm_wg_per_cu();
    UH_LAUNCH(kGemmPersistent[v], dim3(cu_count() * g_gemm_wg[v]), dim3(G_THREADS), 0, s, queue);
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
    const int v = gemm_variant_of(g.precision, cfg);
    if (Batch* b = as_batch(stream)) {
        uh_gemm_args item = g;
        item.precision = g.precision | (cfg << 8);
        record_item(b, FAM_GEMM, &item, sizeof(item), tiles, 1u << v);
        return UH_OK;
    }
    UH_LAUNCH(kGemmPlain[v], dim3(tiles), dim3(G_THREADS), 0, (hipStream_t)stream, g, cfg);
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
    int best = -1;
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
            UH_LAUNCH(kGemmPlain[cfg], dim3(tiles), dim3(G_THREADS), 0, s, g, cfg);
            UH_LAUNCH_CHECK("k_gemm (autotune)");
            UH_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) UH_LAUNCH(kGemmPlain[cfg], dim3(tiles), dim3(G_THREADS), 0, s, g, cfg);
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
    const int rc = run();
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
