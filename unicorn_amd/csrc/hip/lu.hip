// LU kernels for gfx950: the three stages of the reference's blocked LU
// (testSuite/luDecomposition/source/code/src/luDecompositionCUDA.cu:60-193) and the FP64
// iterative-refinement helpers of the solve the config adds.
//
// Conventions follow the reference (luDecomposition.cpp:116-125): column-major, no pivoting;
// step i scales row i right of the diagonal by RN(1/a_ii) and subtracts the rank-1 product of
// column i (unscaled) and row i (scaled). So L is lower with the pivots on its diagonal and U is
// unit upper.
//
// Panel: the reference issues 3 tiny launches per pivot (reciprocal, sscal, sger; 1,533 launches
// per 512 block). Here the block is factored right-looking over 32-column sub-panels, two
// launches each: LU of the 32x32 diagonal sub-block in registers (the reference's step order)
// with the sub-row and sub-column triangular solves spread over workgroups, then the rank-32
// update of the remainder over 64x64 tiles. The block stays in L2 (1 MiB).
// TRSM: a workgroup owns a slab of 32 right-hand-side vectors held in LDS and walks the triangle
// in 32-row blocks: an fp32-MFMA block update from the LDS-staged row panel, then a 32-step
// substitution in registers.
#include "uh_internal.h"

namespace uh {

static constexpr int LU_NB = 32;

// In-register lower solve of 32 unknowns, right-looking: for j = 0..31, x_j is final once
// scaled by scale[j] (when SCALE), then x_i -= C_j[i] x_j for i > j, where column C_j is the 32
// floats at cols + j * ld (16-byte aligned). Every x_i receives its terms in the order j = 0, 1,
// .., i-1, the fma sequence of the row-by-row form. Column j+1 is read from LDS while column j
// is applied. The steps are instantiated one by one (straight-line code: a loop around the
// memory barrier would not be unrolled), and the compiler-only memory barrier after each step
// keeps the loads from all being hoisted to the top, which would need ~600 registers.
template <bool SCALE, int J>
__device__ __forceinline__ void lower_solve32_step(float (&x)[32], const float* cols, int ld, const float (&sc)[32],
                                                   float4 (&cur)[8], float4 (&nxt)[8]) {
    if constexpr (J < 32) {
        if constexpr (J + 1 < 32) {
#pragma unroll
            for (int q = (J + 2) / 4; q < 8; ++q) nxt[q] = *(const float4*)(cols + (J + 1) * ld + q * 4);
        }
        if constexpr (SCALE) x[J] = x[J] * sc[J];
#pragma unroll
        for (int q = (J + 1) / 4; q < 8; ++q) {
            const float tv[4] = {cur[q].x, cur[q].y, cur[q].z, cur[q].w};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (q * 4 + c > J) x[q * 4 + c] = fmaf(-tv[c], x[J], x[q * 4 + c]);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) cur[q] = nxt[q];
        asm volatile("" ::: "memory");
        lower_solve32_step<SCALE, J + 1>(x, cols, ld, sc, cur, nxt);
    }
}

template <bool SCALE>
__device__ __forceinline__ void lower_solve32(float (&x)[32], const float* cols, int ld, const float* scale) {
    float sc[32];
    if constexpr (SCALE) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 t = *(const float4*)(scale + q * 4);
            sc[q * 4] = t.x; sc[q * 4 + 1] = t.y; sc[q * 4 + 2] = t.z; sc[q * 4 + 3] = t.w;
        }
    }
    float4 cur[8], nxt[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        cur[q] = *(const float4*)(cols + q * 4);
        nxt[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    lower_solve32_step<SCALE, 0>(x, cols, ld, sc, cur, nxt);
}

// ---------------------------------------------------------------- panel

// Sub-step kb of the panel, launch 1. Every workgroup loads the nb x nb diagonal sub-block
// (padded to 32 x 32 with the identity) and factors it in registers of wave 0, lane c holding
// column c: step i reads the pivot and column i across lanes, scales row i by RN(1/pivot) and
// applies the rank-1 update, the reference's per-pivot order. Every workgroup computes the same
// bits, so workgroup 0 stores its copy and the others use theirs:
//   workgroups 1 .. nr:       256 columns each of the sub-row panel, U12 = L11^-1 A12;
//   workgroups nr+1 .. 2 nr:  256 rows each of the sub-column panel, L21 = A21 U11^-1;
// one column or row per thread, solved in registers. The panel slice is loaded (coalesced, into
// registers) before the factorisation and passes through LDS on its way in and out.
__global__ void __launch_bounds__(256) k_lu_step(float* a_, uint64_t lda, int n, int kb) {
    __shared__ __attribute__((aligned(16))) float d[LU_NB][36];    // d[r][c]: row r of the factored sub-block
    __shared__ __attribute__((aligned(16))) float dt[LU_NB][36];   // dt[c][r] = d[r][c]
    __shared__ __attribute__((aligned(16))) float rcp[LU_NB];
    __shared__ float px[256][LU_NB + 1];                           // px[t][i]: thread t's column or row
    UH_GLOBAL float* a = as_global(a_);
    const int tid = threadIdx.x;
    const int nb = min(LU_NB, n - kb);
    const int rest = n - kb - nb;
    const int nr = (rest + 255) / 256;
    const int role = blockIdx.x == 0 ? 0 : (blockIdx.x <= (unsigned)nr ? 1 : 2);
    const int base = kb + nb + ((int)blockIdx.x - 1 - (role == 2 ? nr : 0)) * 256;   // first column / row
    const int cnt = min(256, n - base);                                                // columns / rows here
    // element q of this thread's share: role 1 column base + q * 8 + tid / 32, row kb + tid % 32;
    // role 2 row base + tid, column kb + q. Addresses are clamped into the block, values zeroed.
    float ld[LU_NB];
    if (role != 0) {
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) {
            int c, r;
            if (role == 1) { c = base + q * 8 + (tid >> 5); r = kb + (tid & 31); }
            else { c = kb + q; r = base + tid; }
            const bool ok = (role == 1) ? (q * 8 + (tid >> 5) < cnt && (tid & 31) < nb) : (tid < cnt && q < nb);
            const float v = a[(uint64_t)min(c, n - 1) * lda + min(r, n - 1)];
            ld[q] = ok ? v : 0.f;
        }
    }
    if (tid < 64) {
        const int c = tid & 31;
        float col[LU_NB];
#pragma unroll
        for (int r = 0; r < LU_NB; ++r) {
            const float v = a[(uint64_t)(kb + min(c, nb - 1)) * lda + kb + min(r, nb - 1)];
            col[r] = (r < nb && c < nb) ? v : (r == c ? 1.f : 0.f);
        }
#pragma unroll
        for (int i = 0; i < LU_NB; ++i) {
            const float ri = 1.0f / __int_as_float(__builtin_amdgcn_readlane(__float_as_int(col[i]), i));
            const bool act = c > i;
            if (act) col[i] = col[i] * ri;
#pragma unroll
            for (int r = i + 1; r < LU_NB; ++r) {
                const float l = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(col[r]), i));
                if (act) col[r] = fmaf(-l, col[i], col[r]);
            }
            if (tid == 0) rcp[i] = ri;
        }
        if (tid < 32) {
#pragma unroll
            for (int r = 0; r < LU_NB; ++r) {
                d[r][c] = col[r];
                dt[c][r] = col[r];
            }
        }
    }
    if (role != 0) {
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) {
            if (role == 1) px[q * 8 + (tid >> 5)][tid & 31] = ld[q];
            else px[tid][q] = ld[q];
        }
    }
    __syncthreads();
    if (role == 0) {
        for (int e = tid; e < nb * nb; e += 256) {
            const int r = e % nb, c = e / nb;
            a[(uint64_t)(kb + c) * lda + kb + r] = d[r][c];
        }
        return;
    }
    float x[LU_NB];
#pragma unroll
    for (int i = 0; i < LU_NB; ++i) x[i] = px[tid][i];
    if (role == 1)   // U12 column: x_i = (x_i - sum_{j<i} L11(i, j) x_j) * RN(1/L11(i, i)); column j of L11 is dt[j]
        lower_solve32<true>(x, &dt[0][0], 36, rcp);
    else             // L21 row: x_j = x_j - sum_{i<j} x_i U11(i, j), U11 unit upper; row i of U11 is d[i]
        lower_solve32<false>(x, &d[0][0], 36, nullptr);
#pragma unroll
    for (int i = 0; i < LU_NB; ++i) px[tid][i] = x[i];
    __syncthreads();
    if (role == 1) {
        for (int e = tid; e < cnt * 32; e += 256) {
            const int c = e >> 5, i = e & 31;
            if (i < nb) a[(uint64_t)(base + c) * lda + kb + i] = px[c][i];
        }
    } else {
        for (int q = 0; q < nb; ++q)
            if (tid < cnt) a[(uint64_t)(kb + q) * lda + base + tid] = px[tid][q];
    }
}

// Sub-step kb of the panel, launch 2: A22 -= L21 U12 (K = 32) over 64 x 64 tiles; a thread owns
// 4 consecutive rows of 4 columns and accumulates k = 0..31 in order from the element's value.
static constexpr int LU_TL = 68;   // LDS row stride of the 64-wide staged panels (16-byte rows)

__global__ void __launch_bounds__(256) k_lu_rank32(float* a_, uint64_t lda, int n, int kb) {
    __shared__ __attribute__((aligned(16))) float tl[LU_NB][LU_TL];   // tl[k][r] = L21(tr + r, k)
    __shared__ __attribute__((aligned(16))) float tu[LU_NB][LU_TL];   // tu[k][c] = U12(k, tc + c)
    UH_GLOBAL float* a = as_global(a_);
    const int tid = threadIdx.x;
    const int r0 = kb + LU_NB;
    const int tiles = (n - r0 + 63) / 64;
    const int tr = r0 + (int)(blockIdx.x % (unsigned)tiles) * 64;
    const int tc = r0 + (int)(blockIdx.x / (unsigned)tiles) * 64;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + 256 * q;
        const int r = e & 63, k = e >> 6;
        tl[k][r] = tr + r < n ? a[(uint64_t)(kb + k) * lda + tr + r] : 0.f;
        const int k2 = e & 31, c = e >> 5;
        tu[k2][c] = tc + c < n ? a[(uint64_t)(tc + c) * lda + kb + k2] : 0.f;
    }
    __syncthreads();
    const int rq = (tid & 15) * 4, cq = (tid >> 4) * 4;
    float s[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool ok = tr + rq + r < n && tc + cq + c < n;
            s[c][r] = ok ? a[(uint64_t)(tc + cq + c) * lda + tr + rq + r] : 0.f;
        }
#pragma unroll 8
    for (int k = 0; k < LU_NB; ++k) {
        const float4 l = *(const float4*)&tl[k][rq];
        const float4 u = *(const float4*)&tu[k][cq];
        const float lv[4] = {l.x, l.y, l.z, l.w}, uv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) s[c][r] = fmaf(-lv[r], uv[c], s[c][r]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (tr + rq + r < n && tc + cq + c < n) a[(uint64_t)(tc + cq + c) * lda + tr + rq + r] = s[c][r];
}

// ---------------------------------------------------------------- TRSM

// A workgroup solves a slab of 32 right-hand-side vectors in place. Vector v (0..31 of the slab
// starting at s0) has elements x_v[i] = X[elem(v, i)], i = 0..n-1, and the triangle T' is lower:
// x_i = (b_i - sum_{j<i} T'(i,j) x_j) * rdiag_i with
//   kind 0 (L^-1 X, X n x m col-major):      T'(i,j) = L(i,j), rdiag = RN(1/L(i,i));
//   kind 1 (X U^-1, X m x n col-major):      T'(i,j) = U(j,i), rdiag = 1;
//   kind 2 (U^-1 X, U unit upper, X n x m):  the same lower form on reversed indices.
// The slab lives in LDS. Rows go in blocks of 32. For block b the four waves each apply a
// quarter of the block update sum_{j<b} T'(b+i, j) x_j with v_mfma_f32_32x32x2_f32 from the
// row panel T'(b..b+31, 0..b-1) staged in LDS; then wave 0 adds the partials in a fixed order
// and substitutes through the 32 x 32 diagonal block (lane v holding vector v's 32 values in
// registers) while waves 1..3 store the next block's row panel, which they loaded into registers
// one block earlier, and issue the loads of the one after. Two barriers per block.
static constexpr int TRSM_MAX_N = 512;
static constexpr int TS_LD = 33;   // LDS row stride of the slab
static constexpr int TS_PU = 24;   // 16-byte panel units per prefetching thread (192 threads)

typedef float ts_f32x16 __attribute__((ext_vector_type(16)));

struct TrsmSmem {
    float xs[TRSM_MAX_N * TS_LD];        // xs[i * TS_LD + v] = x_v[i]
    float pt[TRSM_MAX_N * 32];           // T'(b + i, j), j < b, at pt_idx(j, i)
    float red[3][16][64];                // partial block updates of waves 1..3 (MFMA register layout)
    __attribute__((aligned(16))) float rd[TRSM_MAX_N];
    __attribute__((aligned(16))) float dg[2][32][36];   // dg[(b / 32) & 1][j][i] = T'(b + i, b + j)
    __attribute__((aligned(16))) float sink[192][4];    // panel units outside the block's columns
};

// Row j of the panel is 32 floats; its 16-byte groups are XOR-swizzled by (j / 4) % 8 so that
// the transposed unit writes of kind 1 spread over the banks. A permutation within the row:
// the MFMA operand reads (two whole rows per instruction) stay conflict-free.
__device__ __forceinline__ int pt_idx(int j, int i) { return j * 32 + (i ^ (((j >> 2) & 7) << 2)); }

__device__ __forceinline__ float trsm_tval(const TrsmItem& it, const UH_GLOBAL float* tri, int i, int j) {
    if (it.kind == 0) return tri[(uint64_t)j * it.ldt + i];
    if (it.kind == 1) return tri[(uint64_t)i * it.ldt + j];
    return tri[(uint64_t)(it.n - 1 - j) * it.ldt + (it.n - 1 - i)];
}

// The row panel of block b, T'(b + i, j) for i < 32 and j < b + 32, moves in 16-byte units; the
// 192 prefetching threads (tp = 0..191) each own TS_PU of them, all with rows 4 (tp % 8) .. +3:
//   kinds 0 and 2 (T' columns contiguous in the triangle): unit q is column j = tp / 8 + 24 q;
//   kind 1 (T' rows contiguous): units 4 q' .. 4 q' + 3 are the 4 rows at columns 4 j4 .. 4 j4 + 3,
//     j4 = tp / 8 + 24 q', transposed in registers into 4 columns before the store.
// With ldt, n and the triangle's address multiples of 4 floats the units are single 16-byte
// buffer loads (32-bit offsets: the host checks the triangle spans less than 2 GiB); otherwise
// each is four 4-byte loads. Elements outside the triangle read as zero.
__device__ __forceinline__ void trsm_panel_load(const TrsmItem& it, int b, int tp, float4 (&pf)[TS_PU]) {
    const int n = it.n, ld = (int)it.ldt, J = min(b + 32, n);
    const int i0 = 4 * (tp & 7), jt = tp >> 3;
    const int extent = (int)(((int64_t)(n - 1) * (int64_t)ld + n) * 4);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)it.tri, 0, extent, 0x00020000);
    const bool vec = ((ld | n) & 3) == 0 && ((uintptr_t)it.tri & 15) == 0;
    auto ld1 = [&](bool ok, int off) -> float {
        const float v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, ok ? (uint32_t)off * 4u : 0u, 0, 0));
        return ok ? v : 0.f;
    };
    auto ld4 = [&](bool ok, int off) -> float4 {
        const float4 v = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, ok ? (uint32_t)off * 4u : 0u, 0, 0));
        return ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (it.kind != 1) {
        const bool rev = it.kind == 2;
#pragma unroll
        for (int q = 0; q < TS_PU; ++q) {
            const int j = jt + 24 * q;
            // element (b + i0 + t, j): kind 0 at j ld + b + i0 + t; kind 2 at (n-1-j) ld + n-1-b-i0-t
            const int o = rev ? (n - 1 - j) * ld + (n - 1 - b - i0) : j * ld + b + i0;
            if (vec) {
                const bool ok = j < J && b + i0 < n;
                const float4 v = ld4(ok, rev ? o - 3 : o);
                pf[q] = rev ? make_float4(v.w, v.z, v.y, v.x) : v;
            } else {
                const int s = rev ? -1 : 1;
                pf[q] = make_float4(ld1(j < J && b + i0 < n, o), ld1(j < J && b + i0 + 1 < n, o + s),
                                    ld1(j < J && b + i0 + 2 < n, o + 2 * s), ld1(j < J && b + i0 + 3 < n, o + 3 * s));
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < TS_PU / 4; ++q) {
            const int j = 4 * (jt + 24 * q);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bool rok = b + i0 + r < n;
                const int o = (b + i0 + r) * ld + j;   // element (b + i0 + r, j .. j + 3)
                if (vec) pf[4 * q + r] = ld4(rok && j < J, o);
                else pf[4 * q + r] = make_float4(ld1(rok && j < J, o), ld1(rok && j + 1 < J, o + 1),
                                                 ld1(rok && j + 2 < J, o + 2), ld1(rok && j + 3 < J, o + 3));
            }
        }
    }
}

// Branch-free: a unit (rows i0 .. i0 + 3 of column j) lands in the panel (j < b), the diagonal
// block or the thread's sink.
__device__ __forceinline__ void trsm_unit_store(TrsmSmem& sm, int b, int j, int i0, int tp, float4 v) {
    float* dg = &sm.dg[(b >> 5) & 1][0][0];
    float* dst = j < b ? &sm.pt[pt_idx(j, i0)] : (j < b + 32 ? &dg[(j - b) * 36 + i0] : &sm.sink[tp][0]);
    *(float4*)dst = v;
}

__device__ __forceinline__ void trsm_panel_store(int kind, int b, int tp, const float4 (&pf)[TS_PU], TrsmSmem& sm) {
    const int i0 = 4 * (tp & 7), jt = tp >> 3;
    if (kind != 1) {
#pragma unroll
        for (int q = 0; q < TS_PU; ++q) trsm_unit_store(sm, b, jt + 24 * q, i0, tp, pf[q]);
    } else {
#pragma unroll
        for (int q = 0; q < TS_PU / 4; ++q) {
            const int j = 4 * (jt + 24 * q);
            const float4 r0 = pf[4 * q], r1 = pf[4 * q + 1], r2 = pf[4 * q + 2], r3 = pf[4 * q + 3];
            trsm_unit_store(sm, b, j, i0, tp, make_float4(r0.x, r1.x, r2.x, r3.x));
            trsm_unit_store(sm, b, j + 1, i0, tp, make_float4(r0.y, r1.y, r2.y, r3.y));
            trsm_unit_store(sm, b, j + 2, i0, tp, make_float4(r0.z, r1.z, r2.z, r3.z));
            trsm_unit_store(sm, b, j + 3, i0, tp, make_float4(r0.w, r1.w, r2.w, r3.w));
        }
    }
}

// Quarter `wave` of the block update of block b: D(i, v) = sum T'(b+i, j) x_j[v] over
// j in [wave b/4, (wave+1) b/4) (zero for b = 0).
__device__ __forceinline__ ts_f32x16 trsm_block_update(const TrsmSmem& sm, int b, int wave, int lane) {
    ts_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const int q = b >> 2;   // a multiple of 8
    const int jo = lane >> 5, c = lane & 31;
    for (int j = wave * q; j < (wave + 1) * q; j += 8) {
        float av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            av[u] = sm.pt[pt_idx(j + 2 * u + jo, c)];
            bv[u] = sm.xs[(j + 2 * u + jo) * TS_LD + c];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], bv[u], acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ void trsm_slab(const TrsmItem& it, uint32_t slab, TrsmSmem& sm) {
    const int n = it.n, m = it.m, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: the role branches are scalar
    const int tp = tid - 64;
    const int np = (n + 31) & ~31;   // rows padded to whole blocks
    const int s0 = (int)slab * 32;
    const int nv = min(32, m - s0);
    UH_GLOBAL float* x = as_global(it.x);
    const UH_GLOBAL float* tri = as_global(it.tri);
    auto xel = [&](int v, int i) -> UH_GLOBAL float* {
        if (it.kind == 0) return x + (uint64_t)(s0 + v) * it.ldx + i;          // column s0+v, row i
        if (it.kind == 1) return x + (uint64_t)i * it.ldx + (s0 + v);          // row s0+v, column i
        return x + (uint64_t)(s0 + v) * it.ldx + (n - 1 - i);                  // column, reversed
    };
    // the slab, 16 loads in flight per thread
    for (int e0 = 0; e0 < np * 32; e0 += 256 * 16) {
        float v16[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = e0 + u * 256 + tid;
            int i, v;
            if (it.kind != 1) { i = e % np; v = e / np; } else { v = e % 32; i = e / 32; }
            const bool ok = e < np * 32 && v < nv && i < n;
            const UH_GLOBAL float* p = ok ? xel(v, i) : x;
            const float t = *p;
            v16[u] = ok ? t : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = e0 + u * 256 + tid;
            int i, v;
            if (it.kind != 1) { i = e % np; v = e / np; } else { v = e % 32; i = e / 32; }
            if (e < np * 32) sm.xs[i * TS_LD + v] = v16[u];
        }
    }
    for (int i = tid; i < np; i += 256) sm.rd[i] = (it.kind == 0 && i < n) ? 1.0f / trsm_tval(it, tri, i, i) : 1.0f;
    // The two roles run separate loops with the same two barriers per block, so the prefetch
    // registers of waves 1..3 do not exist on wave 0's path through the substitution.
    if (wave == 0) {
        __syncthreads();   // slab, rd and block 0's panel in LDS
        for (int b = 0; b < n; b += 32) {
            ts_f32x16 acc = trsm_block_update(sm, b, 0, lane);
            __syncthreads();
            if (b > 0) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    float t = acc[r];
                    t += sm.red[0][r][lane];
                    t += sm.red[1][r][lane];
                    t += sm.red[2][r][lane];
                    const int row = (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                    sm.xs[(b + row) * TS_LD + (lane & 31)] -= t;
                }
            }
            // lanes 32..63 wrote rows that lanes 0..31 read below
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < 32) {
                const int v = lane;
                float xv[32];
#pragma unroll
                for (int i = 0; i < 32; ++i) xv[i] = sm.xs[(b + i) * TS_LD + v];
                lower_solve32<true>(xv, &sm.dg[(b >> 5) & 1][0][0], 36, &sm.rd[b]);
#pragma unroll
                for (int i = 0; i < 32; ++i) sm.xs[(b + i) * TS_LD + v] = xv[i];
            }
            __syncthreads();
        }
    } else {
        float4 pf[TS_PU];
        trsm_panel_load(it, 0, tp, pf);
        trsm_panel_store(it.kind, 0, tp, pf, sm);
        if (32 < n) trsm_panel_load(it, 32, tp, pf);
        __syncthreads();
        for (int b = 0; b < n; b += 32) {
            const ts_f32x16 acc = trsm_block_update(sm, b, wave, lane);
            if (b > 0) {
#pragma unroll
                for (int r = 0; r < 16; ++r) sm.red[wave - 1][r][lane] = acc[r];
            }
            __syncthreads();
            if (b + 32 < n) {
                trsm_panel_store(it.kind, b + 32, tp, pf, sm);
                if (b + 64 < n) trsm_panel_load(it, b + 64, tp, pf);
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < n * 32; e += 256) {
        int i, v;
        if (it.kind != 1) { i = e % n; v = e / n; } else { v = e % 32; i = e / 32; }
        if (v < nv) *xel(v, i) = sm.xs[i * TS_LD + v];
    }
}

__global__ void __launch_bounds__(256) k_trsm(TrsmItem it) {
    __shared__ TrsmSmem sm;
    trsm_slab(it, blockIdx.x, sm);
}

__global__ void __launch_bounds__(256) k_trsm_persistent(const uint8_t* queue) {
    __shared__ TrsmSmem sm;
    __shared__ uint32_t s_tile;
    QueueHeader* h = (QueueHeader*)queue;
    const TrsmItem* items = (const TrsmItem*)(queue + h->items_offset);
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
        trsm_slab(uniform_load(items + item), local, sm);
        __syncthreads();
    }
}

uint32_t trsm_tiles(const TrsmItem& a) { return (uint32_t)((a.m + 31) / 32); }

int trsm_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_trsm_persistent, dim3(cu_count()), dim3(256), 0, s, queue);
    UH_LAUNCH_CHECK("k_trsm_persistent");
    g_persistent_launches++;
    return UH_OK;
}

// ---------------------------------------------------------------- blocked triangular solve

// The refinement's solve L U z = r for nrhs <= 8 vectors: one launch inverting the 128 x 128
// diagonal blocks of both triangles, then one launch per triangle. Rows are cut into 128-row
// blocks and a workgroup claims the next block by ticket (forward: top down; backward: bottom
// up), so a block depends only on blocks claimed earlier, by workgroups that are already
// running: the lowest unfinished ticket can always proceed, whatever the number of resident
// workgroups. A workgroup streams its row panel one 128 x 128 tile ahead (those loads depend on
// no other block), waits for each tile's block flag before reading that block's solution,
// multiplies its remaining right-hand side by the inverted diagonal block and publishes its
// block: stores, agent-scope release, flag. The dependency chain is N/128 blocks long and its
// diagonal step is a parallel product, not a 128-step substitution. Every wait is bounded: past
// the limit the block is written as NaN and published anyway, so every workgroup exits and the
// refinement sees a non-finite residual.
static constexpr int TV_B = 128;          // rows per block
static constexpr int TV_H = TV_B / 2;     // columns of a tile per thread (two column halves)

struct TrsvArgs {
    const float* lu;
    uint64_t ld;
    float* z;          // n x nrhs column-major, leading dimension n, solved in place
    const float* inv;  // inverted diagonal blocks, [block][triangle] 128 x 128 column-major
    uint32_t* sync;    // [0] ticket, [1 + t] flag of ticket t; zero before the launch
    int32_t n, nrhs, upper, spin_limit;
};

// Inverse of diagonal block p of the lower non-unit (blockIdx.y = 0) or the upper unit (1)
// triangle of the packed factors; rows and columns past n are taken as the identity. Thread j
// computes column j, row by row: every lane runs the same trip count (the terms against the
// zero part of its column add nothing), reads row i of the block as 4-wide broadcasts and keeps
// four independent partial sums.
__global__ void __launch_bounds__(TV_B) k_trsv_diag_inv(const float* lu, uint64_t ld, int n, float* inv) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    __shared__ __attribute__((aligned(16))) float T[TV_B][TV_B + 4];
    __shared__ float X[TV_B][TV_B + 1];
    const int p = blockIdx.x, upper = blockIdx.y, j = threadIdx.x;
    const int b0 = p * TV_B;
    const UH_GLOBAL float* A = as_global(lu);
    for (int c = 0; c < TV_B; ++c) {          // column c, row j: coalesced over the threads
        const int gi = b0 + j, gc = b0 + c;
        const bool in = gi < n && gc < n && (upper ? c > j : c <= j);
        const UH_GLOBAL float* q = in ? A + (uint64_t)gc * ld + gi : A;
        const float v = *q;
        T[j][c] = in ? v : (c == j ? 1.f : 0.f);
    }
    __syncthreads();
    auto dot = [&](int i, int k0, int k1) {   // sum over k in [k0, k1) of T[i][k] X[k][j], k0 % 4 == 0
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        int k = k0;
        for (; k + 4 <= k1; k += 4) {
            const f32x4 t = *(const f32x4*)&T[i][k];
            s0 = fmaf(t.x, X[k][j], s0);
            s1 = fmaf(t.y, X[k + 1][j], s1);
            s2 = fmaf(t.z, X[k + 2][j], s2);
            s3 = fmaf(t.w, X[k + 3][j], s3);
        }
        for (; k < k1; ++k) s0 = fmaf(T[i][k], X[k][j], s0);
        return (s0 + s1) + (s2 + s3);
    };
    if (!upper) {
        for (int i = 0; i < TV_B; ++i) {      // L X = I, rows top down
            const float s = (i == j ? 1.f : 0.f) - dot(i, 0, i);
            X[i][j] = s / T[i][i];
        }
    } else {
        for (int i = TV_B - 1; i >= 0; --i) { // U X = I, rows bottom up (unit diagonal)
            const int k0 = (i + 1 + 3) & ~3;  // the terms below k0 are summed one by one
            float s = 0.f;
            for (int k = i + 1; k < k0 && k < TV_B; ++k) s = fmaf(T[i][k], X[k][j], s);
            X[i][j] = (i == j ? 1.f : 0.f) - (s + dot(i, k0 < TV_B ? k0 : TV_B, TV_B));
        }
    }
    __syncthreads();
    UH_GLOBAL float* out = as_global(inv) + ((size_t)p * 2 + upper) * TV_B * TV_B;
    for (int c = 0; c < TV_B; ++c) out[c * TV_B + j] = X[j][c];
}

__global__ void __launch_bounds__(256) k_trsv_blocked(TrsvArgs a) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    __shared__ __attribute__((aligned(16))) float xs[2][TV_B][8];   // a column block's solution
    __shared__ __attribute__((aligned(16))) float red[2][TV_B][8];  // the column halves' sums
    __shared__ __attribute__((aligned(16))) float rs[TV_B][8];      // right-hand side of the diagonal step
    __shared__ uint32_t s_t;
    __shared__ int s_fail;
    const int tid = threadIdx.x, i = tid & (TV_B - 1), h = tid >> 7;
    const int n = a.n, nb = (n + TV_B - 1) / TV_B;
    const int64_t ld = (int64_t)a.ld;
    UH_GLOBAL float* z = as_global(a.z);
    for (;;) {
        if (tid == 0) {
            s_t = atomicAdd(a.sync, 1u);
            s_fail = 0;
        }
        __syncthreads();
        const int t = (int)__builtin_amdgcn_readfirstlane(s_t);
        if (t >= nb) break;
        const int p = a.upper ? nb - 1 - t : t;   // row block
        const int R = p * TV_B + i;                // this thread's row
        auto colblock = [&](int d) { return a.upper ? nb - 1 - d : d; };   // row block of ticket d
        // Tile loads through a buffer resource based at the tile's first column: the lane offset
        // (row) is fixed and the column step is a scalar offset, so the 64 loads need no address
        // registers. Columns past n fall outside the resource and read 0; rows past n are masked.
        const uint32_t row_off = (uint32_t)R * 4u;
        const uint32_t col_step = (uint32_t)__builtin_amdgcn_readfirstlane((int)(ld * 4));
        auto load_tile = [&](int d, float (&m)[TV_H]) {   // m[c] = M(R, 128 colblock(d) + 64 h + c)
            const int c0 = __builtin_amdgcn_readfirstlane(colblock(d) * TV_B + h * TV_H);
            // bytes from column c0 to the end of the last column of the 64 that exist (none: 0)
            const int cols = min(n - c0, TV_H);
            const int bytes = cols > 0 ? (int)((uint32_t)(cols - 1) * col_step + (uint32_t)n * 4u) : 0;
            const __amdgpu_buffer_rsrc_t rs_m = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(a.lu + (int64_t)c0 * ld), 0, bytes, 0x00020000);
            // rows past n read finite values that only reach rows whose results are dropped
#pragma unroll
            for (int c = 0; c < TV_H; ++c)
                m[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_m, row_off, c * col_step, 0));
        };
        // row i of the inverted diagonal block, this half's columns
        float dv[TV_H];
        {
            const __amdgpu_buffer_rsrc_t rs_i = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(a.inv + ((int64_t)p * 2 + a.upper) * TV_B * TV_B + (int64_t)h * TV_H * TV_B), 0,
                TV_H * TV_B * 4, 0x00020000);
#pragma unroll
            for (int c = 0; c < TV_H; ++c)
                dv[c] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_i, (uint32_t)i * 4u, c * TV_B * 4, 0));
        }
        // this block's own right-hand side (thread (row i, vectors 4h .. 4h+3)), loaded before
        // any wait: it depends on no other block
        float bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = h * 4 + e;
            const bool ok = R < n && k < a.nrhs;
            const UH_GLOBAL float* q = ok ? z + (int64_t)k * n + R : z;
            const float b = *q;
            bv[e] = ok ? b : 0.f;
        }
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        // tile d: wait for its block, stage that block's solution, accumulate M(R, .) x
        auto step = [&](int d, const float (&m)[TV_H]) {
            if (tid == 0) {
                uint32_t* f = a.sync + 1 + d;
                int spins = 0;
                while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                    if (++spins > a.spin_limit) {
                        s_fail = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int buf = d & 1;
            {   // xs[buf][c][4 kg .. 4 kg + 3] = z(128 colblock(d) + c, k): thread (c = i, kg = h)
                const int col = colblock(d) * TV_B + i;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int k = h * 4 + e;
                    const bool ok = k < a.nrhs && col < n;
                    const UH_GLOBAL float* q = ok ? z + (int64_t)k * n + col : z;
                    const float x = *q;
                    v[e] = ok ? x : 0.f;
                }
                *(f32x4*)&xs[buf][i][h * 4] = f32x4{v[0], v[1], v[2], v[3]};
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < TV_H; ++c) {
                const f32x4 x0 = *(const f32x4*)&xs[buf][h * TV_H + c][0];
                const f32x4 x1 = *(const f32x4*)&xs[buf][h * TV_H + c][4];
                acc[0] = fmaf(m[c], x0.x, acc[0]);
                acc[1] = fmaf(m[c], x0.y, acc[1]);
                acc[2] = fmaf(m[c], x0.z, acc[2]);
                acc[3] = fmaf(m[c], x0.w, acc[3]);
                acc[4] = fmaf(m[c], x1.x, acc[4]);
                acc[5] = fmaf(m[c], x1.y, acc[5]);
                acc[6] = fmaf(m[c], x1.z, acc[6]);
                acc[7] = fmaf(m[c], x1.w, acc[7]);
            }
        };
        // two tile buffers in turn, the next tile's loads in flight during a step
        float ma[TV_H], mb[TV_H];
        load_tile(0, ma);
        for (int d = 0; d < t; d += 2) {
            if (d + 1 < t) load_tile(d + 1, mb);
            step(d, ma);
            if (d + 1 >= t) break;
            if (d + 2 < t) load_tile(d + 2, ma);
            step(d + 1, mb);
        }
        *(f32x4*)&red[h][i][0] = f32x4{acc[0], acc[1], acc[2], acc[3]};
        *(f32x4*)&red[h][i][4] = f32x4{acc[4], acc[5], acc[6], acc[7]};
        __syncthreads();
        {   // thread (row i, vectors 4h .. 4h+3): rhs minus the two halves, in a fixed order
            const f32x4 s0 = *(const f32x4*)&red[0][i][h * 4];
            const f32x4 s1 = *(const f32x4*)&red[1][i][h * 4];
            const float sum[4] = {s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w};
            float r[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) r[e] = R < n ? bv[e] - sum[e] : 0.f;
            *(f32x4*)&rs[i][h * 4] = f32x4{r[0], r[1], r[2], r[3]};
        }
        __syncthreads();
        float y[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < TV_H; ++c) {
            const f32x4 r0 = *(const f32x4*)&rs[h * TV_H + c][0];
            const f32x4 r1 = *(const f32x4*)&rs[h * TV_H + c][4];
            y[0] = fmaf(dv[c], r0.x, y[0]);
            y[1] = fmaf(dv[c], r0.y, y[1]);
            y[2] = fmaf(dv[c], r0.z, y[2]);
            y[3] = fmaf(dv[c], r0.w, y[3]);
            y[4] = fmaf(dv[c], r1.x, y[4]);
            y[5] = fmaf(dv[c], r1.y, y[5]);
            y[6] = fmaf(dv[c], r1.z, y[6]);
            y[7] = fmaf(dv[c], r1.w, y[7]);
        }
        *(f32x4*)&red[h][i][0] = f32x4{y[0], y[1], y[2], y[3]};   // red was last read before the barrier above
        *(f32x4*)&red[h][i][4] = f32x4{y[4], y[5], y[6], y[7]};
        __syncthreads();
        {
            const f32x4 s0 = *(const f32x4*)&red[0][i][h * 4];
            const f32x4 s1 = *(const f32x4*)&red[1][i][h * 4];
            const float x[4] = {s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w};
            const bool nan = s_fail != 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = h * 4 + e;
                if (R < n && k < a.nrhs) z[(int64_t)k * n + R] = nan ? __int_as_float(0x7fc00000) : x[e];
            }
        }
        // The barrier orders every thread's stores before thread 0's flag store, whose
        // agent-scope release makes them visible to the other XCDs (release is cumulative).
        __syncthreads();
        if (tid == 0) __hip_atomic_store(a.sync + 1 + t, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- refinement helpers

// part[c][i][k] = sum_{j in chunk c} A(i, j) * x(j, k), A fp32 column-major, x fp64; nrhs <= 8
static constexpr int RES_CHUNK = 512;

template <typename XT>
__global__ void __launch_bounds__(256) k_colblock_gemv(const float* a, uint64_t lda, int n_rows, int j0,
                                                        int j1, const XT* x, uint64_t ldx, int nrhs,
                                                        double* part) {
    __shared__ double xs[RES_CHUNK][8];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    const int jb = j0 + c * RES_CHUNK;
    const int je = min(j1, jb + RES_CHUNK);
    for (int e = threadIdx.x; e < (je - jb) * 8; e += 256) {
        const int j = e / 8, k = e % 8;
        xs[j][k] = k < nrhs ? (double)x[(uint64_t)k * ldx + jb + j] : 0.0;
    }
    __syncthreads();
    if (i >= n_rows) return;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = jb; j < je; ++j) {
        const double av = (double)a[(uint64_t)j * lda + i];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = fma(av, xs[j - jb][k], s[k]);
    }
    for (int k = 0; k < nrhs; ++k) part[((uint64_t)c * n_rows + i) * 8 + k] = s[k];
}

// r = b - sum_c part[c]   (fixed chunk order)
__global__ void __launch_bounds__(256) k_residual_sum(const double* part, int chunks, int n, const double* b,
                                                       int nrhs, double* r, float* rf) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < nrhs; ++k) {
        double s = 0.0;
        for (int c = 0; c < chunks; ++c) s += part[((uint64_t)c * n + i) * 8 + k];
        const double v = b[(uint64_t)k * n + i] - s;
        r[(uint64_t)k * n + i] = v;
        if (rf) rf[(uint64_t)k * n + i] = (float)v;
    }
}

__global__ void __launch_bounds__(256) k_refine_update(double* x, const float* z, int n, int nrhs) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < (uint64_t)n * nrhs) x[e] += (double)z[e];
}

__global__ void __launch_bounds__(256) k_colmax_abs(const double* v, int n, int ncols, double* out) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(v[(uint64_t)c * n + i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

}  // namespace uh

using namespace uh;

extern "C" int uh_lu_panel(float* a, uint64_t lda, int n, void* stream) {
    if (!a || n <= 0 || lda < (uint64_t)n) return fail(UH_ERR_ARG, "uh_lu_panel: bad arguments");
    hipStream_t s;
    if (Batch* b = as_batch(stream)) {
        // a panel is a barrier: run what was recorded, then the panel on the batch's stream
        int rc = uh_batch_flush(stream, 0);
        if (rc) return rc;
        s = b->stream;
    } else {
        s = (hipStream_t)stream;
    }
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int rest = n - kb - min(LU_NB, n - kb);
        UH_LAUNCH(k_lu_step, dim3(1 + 2 * ((rest + 255) / 256)), dim3(256), 0, s, a, lda, n, kb);
        UH_LAUNCH_CHECK("k_lu_step");
        g_plain_launches++;
        if (rest > 0) {
            const unsigned tiles = (unsigned)((rest + 63) / 64);
            UH_LAUNCH(k_lu_rank32, dim3(tiles * tiles), dim3(256), 0, s, a, lda, n, kb);
            UH_LAUNCH_CHECK("k_lu_rank32");
            g_plain_launches++;
        }
    }
    return UH_OK;
}

static int trsm_common(const float* t, uint64_t ldt, float* x, uint64_t ldx, int n, int m, int kind, void* stream) {
    if (!t || !x || n <= 0 || m < 0) return fail(UH_ERR_ARG, "uh_trsm: bad arguments");
    if (n > TRSM_MAX_N) return fail(UH_ERR_ARG, "uh_trsm: n=%d above %d", n, TRSM_MAX_N);
    if (ldt < (uint64_t)n) return fail(UH_ERR_ARG, "uh_trsm: ldt too small");
    if (((uint64_t)(n - 1) * ldt + n) * 4 >= (1ull << 31))
        return fail(UH_ERR_ARG, "uh_trsm: the triangle spans 2 GiB or more (n=%d, ldt=%llu)", n,
                    (unsigned long long)ldt);
    if (kind == 0 && ldx < (uint64_t)n) return fail(UH_ERR_ARG, "uh_trsm: ldx too small");
    if (kind == 1 && ldx < (uint64_t)m) return fail(UH_ERR_ARG, "uh_trsm: ldx too small");
    if (m == 0) return UH_OK;
    TrsmItem it{t, x, ldt, ldx, n, m, kind, 0};
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_TRSM, &it, sizeof(it), trsm_tiles(it));
        return UH_OK;
    }
    UH_LAUNCH(k_trsm, dim3(trsm_tiles(it)), dim3(256), 0, (hipStream_t)stream, it);
    UH_LAUNCH_CHECK("k_trsm");
    g_plain_launches++;
    return UH_OK;
}

extern "C" int uh_trsm_lln(const float* l, uint64_t ldl, float* x, uint64_t ldx, int n, int m, void* stream) {
    return trsm_common(l, ldl, x, ldx, n, m, 0, stream);
}

extern "C" int uh_trsm_run(const float* u, uint64_t ldu, float* x, uint64_t ldx, int m, int n, void* stream) {
    return trsm_common(u, ldu, x, ldx, n, m, 1, stream);
}

static hipStream_t plain_stream(void* stream, int* rc) {
    *rc = UH_OK;
    if (Batch* b = as_batch(stream)) {
        *rc = uh_batch_flush(stream, 0);
        return b->stream;
    }
    return (hipStream_t)stream;
}

static double* g_part = nullptr;
static size_t g_part_bytes = 0;
static std::mutex g_part_mtx;

extern "C" int uh_residual_f64(const float* a, uint64_t lda, int n, const double* x, const double* b,
                               int nrhs, double* r, float* rf, void* stream) {
    if (!a || !x || !b || !r || n <= 0 || nrhs <= 0 || nrhs > 8 || lda < (uint64_t)n)
        return fail(UH_ERR_ARG, "uh_residual_f64: bad arguments");
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    const int chunks = (n + RES_CHUNK - 1) / RES_CHUNK;
    const size_t need = (size_t)chunks * n * 8 * sizeof(double);
    std::lock_guard<std::mutex> lk(g_part_mtx);
    if (need > g_part_bytes) {
        if (g_part) { UH_HIP(hipStreamSynchronize(s)); UH_HIP(hipFree(g_part)); g_part = nullptr; g_part_bytes = 0; }
        UH_HIP(hipMalloc(&g_part, need));
        g_part_bytes = need;
    }
    UH_LAUNCH(k_colblock_gemv<double>, dim3((n + 255) / 256, chunks), dim3(256), 0, s, a, lda, n, 0, n,
                       x, (uint64_t)n, nrhs, g_part);
    UH_LAUNCH_CHECK("k_colblock_gemv");
    UH_LAUNCH(k_residual_sum, dim3((n + 255) / 256), dim3(256), 0, s, g_part, chunks, n, b, nrhs, r, rf);
    UH_LAUNCH_CHECK("k_residual_sum");
    g_plain_launches += 2;
    return UH_OK;
}

// Per-stream workspace of the triangular solves (sync words, inverted diagonal blocks): solves
// on one stream run in stream order and share it; solves on different streams get their own.
struct TrsvWorkspace {
    uint32_t* sync = nullptr;
    size_t words = 0;
    float* inv = nullptr;
    size_t inv_floats = 0;
};
static std::map<hipStream_t, TrsvWorkspace> g_trsv_ws;
static std::mutex g_trsv_mtx;
static constexpr int TV_SPIN_LIMIT = 1 << 22;   // polls of one flag, about a second

extern "C" int uh_lu_solve_f32(const float* lu, uint64_t ld, int n, float* rhs, int nrhs, void* stream) {
    if (!lu || !rhs || n <= 0 || nrhs <= 0 || nrhs > 8 || ld < (uint64_t)n)
        return fail(UH_ERR_ARG, "uh_lu_solve_f32: bad arguments");
    if (ld * 4 * TV_H + (uint64_t)n * 4 >= (1ull << 31))   // 64 columns of a tile under 2 GiB
        return fail(UH_ERR_ARG, "uh_lu_solve_f32: leading dimension too large");
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    const int nb = (n + TV_B - 1) / TV_B;
    const size_t words = 2 * (size_t)(nb + 1);   // forward then backward
    const size_t inv_floats = (size_t)nb * 2 * TV_B * TV_B;
    uint32_t* sync;
    float* inv;
    {
        std::lock_guard<std::mutex> lk(g_trsv_mtx);
        auto& e = g_trsv_ws[s];
        if (e.words < words || e.inv_floats < inv_floats) {
            UH_HIP(hipStreamSynchronize(s));
            if (e.sync) UH_HIP(hipFree(e.sync));
            if (e.inv) UH_HIP(hipFree(e.inv));
            e = TrsvWorkspace{};
            UH_HIP(hipMalloc(&e.sync, words * sizeof(uint32_t)));
            e.words = words;
            UH_HIP(hipMalloc(&e.inv, inv_floats * sizeof(float)));
            e.inv_floats = inv_floats;
        }
        sync = e.sync;
        inv = e.inv;
    }
    UH_LAUNCH(k_trsv_diag_inv, dim3((unsigned)nb, 2), dim3(TV_B), 0, s, lu, ld, n, inv);
    UH_LAUNCH_CHECK("k_trsv_diag_inv");
    rc = uh_fill_u32(sync, 0u, words, s);
    if (rc) return rc;
    for (int upper = 0; upper < 2; ++upper) {
        TrsvArgs a{lu, ld, rhs, inv, sync + upper * (nb + 1), n, nrhs, upper, TV_SPIN_LIMIT};
        UH_LAUNCH(k_trsv_blocked, dim3((unsigned)nb), dim3(256), 0, s, a);
        UH_LAUNCH_CHECK("k_trsv_blocked");
    }
    g_plain_launches += 3;
    return UH_OK;
}

__global__ void __launch_bounds__(256) k_refine_init(const double* b, double* x, double* r, float* rf, uint64_t tot) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= tot) return;
    const double v = b[e];
    x[e] = 0.0;
    r[e] = v;
    rf[e] = (float)v;
}

extern "C" int uh_refine_init(const double* b, double* x, double* r, float* rf, int n, int nrhs, void* stream) {
    if (!b || !x || !r || !rf || n < 0 || nrhs < 0) return fail(UH_ERR_ARG, "uh_refine_init: bad arguments");
    const uint64_t tot = (uint64_t)n * nrhs;
    if (!tot) return UH_OK;
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_refine_init, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, b, x, r, rf, tot);
    UH_LAUNCH_CHECK("k_refine_init");
    g_plain_launches++;
    return UH_OK;
}

extern "C" int uh_refine_update(double* x, const float* z, int n, int nrhs, void* stream) {
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    const uint64_t tot = (uint64_t)n * nrhs;
    UH_LAUNCH(k_refine_update, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, x, z, n, nrhs);
    UH_LAUNCH_CHECK("k_refine_update");
    g_plain_launches++;
    return UH_OK;
}

extern "C" int uh_colmax_abs_f64(const double* v, int n, int ncols, double* out, void* stream) {
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_colmax_abs, dim3(ncols), dim3(256), 0, s, v, n, ncols, out);
    UH_LAUNCH_CHECK("k_colmax_abs");
    g_plain_launches++;
    return UH_OK;
}
