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
// per 512 block). Here one 1024-thread workgroup factors the block with a blocked right-looking
// scheme over 32-column sub-panels: LU of the 32x32 diagonal sub-block in LDS (the reference's
// step order), triangular solves for the sub-row and sub-column panels, and a rank-32 update of
// the remainder through LDS tiles. The block stays in L2 (1 MiB).
// TRSM: a workgroup owns a slab of 32 right-hand-side vectors held in LDS and walks the triangle
// in 32x32 blocks: a block-dot update from LDS-staged triangle tiles, then a 32-step substitution.
#include "uh_internal.h"

namespace uh {

static constexpr int LU_NB = 32;

// ---------------------------------------------------------------- panel

__global__ void __launch_bounds__(1024) k_lu_panel(float* a, uint64_t lda, int n) {
    __shared__ float d[LU_NB][LU_NB + 1];          // diagonal sub-block
    __shared__ float tl[64][LU_NB + 1];            // L21 tile (rows x 32)
    __shared__ float tu[LU_NB][64 + 1];            // U12 tile (32 x cols)
    __shared__ float rcp[LU_NB];
    const int tid = threadIdx.x;
    for (int kb = 0; kb < n; kb += LU_NB) {
        const int nb = min(LU_NB, n - kb);
        // 1. diagonal sub-block: load, factor (reference step order), store
        for (int e = tid; e < nb * nb; e += 1024) {
            const int r = e % nb, c = e / nb;
            d[r][c] = a[(uint64_t)(kb + c) * lda + kb + r];
        }
        __syncthreads();
        for (int i = 0; i < nb; ++i) {
            if (tid == 0) rcp[i] = 1.0f / d[i][i];
            __syncthreads();
            const float ri = rcp[i];
            for (int c = i + 1 + tid; c < nb; c += 1024) d[i][c] = d[i][c] * ri;
            __syncthreads();
            const int m = nb - i - 1;
            for (int e = tid; e < m * m; e += 1024) {
                const int r = i + 1 + e % m, c = i + 1 + e / m;
                d[r][c] = fmaf(-d[r][i], d[i][c], d[r][c]);
            }
            __syncthreads();
        }
        for (int e = tid; e < nb * nb; e += 1024) {
            const int r = e % nb, c = e / nb;
            a[(uint64_t)(kb + c) * lda + kb + r] = d[r][c];
        }
        const int rest = n - kb - nb;
        if (rest <= 0) break;
        // 2. sub-row panel: U12 = L11^-1 A12 (one thread per column, multiply by RN(1/l_ii))
        for (int c = kb + nb + tid; c < n; c += 1024) {
            float* col = a + (uint64_t)c * lda + kb;
            float x[LU_NB];
            for (int i = 0; i < nb; ++i) x[i] = col[i];
            for (int i = 0; i < nb; ++i) {
                float s = x[i];
                for (int j = 0; j < i; ++j) s = fmaf(-d[i][j], x[j], s);
                x[i] = s * rcp[i];
            }
            for (int i = 0; i < nb; ++i) col[i] = x[i];
        }
        // 3. sub-column panel: L21 = A21 U11^-1, U11 unit upper (one thread per row)
        for (int r = kb + nb + tid; r < n; r += 1024) {
            float x[LU_NB];
            for (int j = 0; j < nb; ++j) x[j] = a[(uint64_t)(kb + j) * lda + r];
            for (int j = 0; j < nb; ++j) {
                float s = x[j];
                for (int i = 0; i < j; ++i) s = fmaf(-x[i], d[i][j], s);
                x[j] = s;
            }
            for (int j = 0; j < nb; ++j) a[(uint64_t)(kb + j) * lda + r] = x[j];
        }
        __syncthreads();
        // 4. trailing update A22 -= L21 U12 in 64x64 tiles (4 outputs per thread)
        const int r0 = kb + nb;
        const int tiles = (rest + 63) / 64;
        for (int t = 0; t < tiles * tiles; ++t) {
            const int tr = r0 + (t % tiles) * 64, tc = r0 + (t / tiles) * 64;
            for (int e = tid; e < 64 * nb; e += 1024) {
                const int r = e % 64, k = e / 64;
                tl[r][k] = (tr + r < n) ? a[(uint64_t)(kb + k) * lda + tr + r] : 0.f;
                const int c = e % 64, k2 = e / 64;
                tu[k2][c] = (tc + c < n) ? a[(uint64_t)(tc + c) * lda + kb + k2] : 0.f;
            }
            __syncthreads();
            const int r = tid % 64, cq = (tid / 64) * 4;
            if (tr + r < n) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int c = cq + q;
                    if (tc + c >= n) continue;
                    float* p = a + (uint64_t)(tc + c) * lda + tr + r;
                    float s = *p;
                    for (int k = 0; k < nb; ++k) s = fmaf(-tl[r][k], tu[k][c], s);
                    *p = s;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- TRSM

// Solves 32 vectors in place. Vector v (0..31 of the slab starting at s0) has elements
// x_v[i] = X[elem(v, i)], i = 0..n-1. The triangle T' is lower: x_i = (b_i - sum_{j<i} T'(i,j) x_j)
// * rdiag_i with T'(i,j) = L(i,j), rdiag = RN(1/L(i,i)) for kind 0 (L^-1 X, X n x m col-major);
// T'(i,j) = U(j,i), rdiag = 1 for kind 1 (X U^-1, X m x n col-major); and for kind 2 (U^-1 X,
// U unit upper, X n x m col-major) the same lower form on reversed indices.
__device__ __forceinline__ void trsm_slab(const TrsmItem& it, uint32_t slab, float* xs /* n x 33 */,
                                          float (*tt)[33], float* rd) {
    const int n = it.n, m = it.m, tid = threadIdx.x;
    const int s0 = (int)slab * 32;
    const int nv = min(32, m - s0);
    UH_GLOBAL float* x = as_global(it.x);
    const UH_GLOBAL float* tri = as_global(it.tri);
    auto xel = [&](int v, int i) -> UH_GLOBAL float* {
        if (it.kind == 0) return x + (uint64_t)(s0 + v) * it.ldx + i;          // column s0+v, row i
        if (it.kind == 1) return x + (uint64_t)i * it.ldx + (s0 + v);          // row s0+v, column i
        return x + (uint64_t)(s0 + v) * it.ldx + (n - 1 - i);                  // column, reversed
    };
    auto tval = [&](int i, int j) -> float {
        if (it.kind == 0) return tri[(uint64_t)j * it.ldt + i];
        if (it.kind == 1) return tri[(uint64_t)i * it.ldt + j];
        return tri[(uint64_t)(n - 1 - j) * it.ldt + (n - 1 - i)];
    };
    // load the slab: xs[i * 33 + v]
    for (int e = tid; e < n * 32; e += 256) {
        int i, v;
        if (it.kind != 1) { i = e % n; v = e / n; } else { v = e % 32; i = e / 32; }
        xs[i * 33 + v] = v < nv ? *xel(v, i) : 0.f;
    }
    for (int i = tid; i < n; i += 256) rd[i] = it.kind == 0 ? 1.0f / tval(i, i) : 1.0f;
    __syncthreads();
    for (int b = 0; b < n; b += 32) {
        const int bn = min(32, n - b);
        // block-dot update: x[b+i] -= sum_{j<b} T'(b+i, j) x[j]
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        const int ui = tid % 32, uv = (tid / 32) * 4;   // row ui of the block, vectors uv..uv+3
        for (int jb = 0; jb < b; jb += 32) {
            for (int e = tid; e < 32 * 32; e += 256) {
                const int i = e % 32, j = e / 32;
                tt[i][j] = (i < bn) ? tval(b + i, jb + j) : 0.f;
            }
            __syncthreads();
#pragma unroll 8
            for (int j = 0; j < 32; ++j) {
                const float tv = tt[ui][j];
                const float* xr = xs + (jb + j) * 33 + uv;
                acc[0] = fmaf(tv, xr[0], acc[0]);
                acc[1] = fmaf(tv, xr[1], acc[1]);
                acc[2] = fmaf(tv, xr[2], acc[2]);
                acc[3] = fmaf(tv, xr[3], acc[3]);
            }
            __syncthreads();
        }
        if (ui < bn)
            for (int q = 0; q < 4; ++q) xs[(b + ui) * 33 + uv + q] -= acc[q];
        // diagonal block: stage T'(b+i, b+j) and substitute (one lane per vector)
        for (int e = tid; e < 32 * 32; e += 256) {
            const int i = e % 32, j = e / 32;
            tt[i][j] = (i < bn && j < bn) ? tval(b + i, b + j) : 0.f;
        }
        __syncthreads();
        if (tid < 32) {
            const int v = tid;
            for (int i = 0; i < bn; ++i) {
                float s = xs[(b + i) * 33 + v];
                for (int j = 0; j < i; ++j) s = fmaf(-tt[i][j], xs[(b + j) * 33 + v], s);
                xs[(b + i) * 33 + v] = s * rd[b + i];
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * 32; e += 256) {
        int i, v;
        if (it.kind != 1) { i = e % n; v = e / n; } else { v = e % 32; i = e / 32; }
        if (v < nv) *xel(v, i) = xs[i * 33 + v];
    }
}

static constexpr int TRSM_MAX_N = 512;

__global__ void __launch_bounds__(256) k_trsm(TrsmItem it) {
    extern __shared__ float xs[];
    __shared__ float tt[32][33];
    __shared__ float rd[TRSM_MAX_N];
    trsm_slab(it, blockIdx.x, xs, tt, rd);
}

__global__ void __launch_bounds__(256) k_trsm_persistent(const uint8_t* queue) {
    __shared__ float xs[TRSM_MAX_N * 33];
    __shared__ float tt[32][33];
    __shared__ float rd[TRSM_MAX_N];
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
        trsm_slab(uniform_load(items + item), local, xs, tt, rd);
        __syncthreads();
    }
}

uint32_t trsm_tiles(const TrsmItem& a) { return (uint32_t)((a.m + 31) / 32); }

int trsm_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_trsm_persistent, dim3(cu_count() * 2), dim3(256), 0, s, queue);
    UH_LAUNCH_CHECK("k_trsm_persistent");
    g_persistent_launches++;
    return UH_OK;
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

// y(i, k) -= sum_{j in [j0, j1)} A(i, j) * z(j, k) for rows [i0, i1): fp32
__global__ void __launch_bounds__(256) k_update_f32(const float* a, uint64_t lda, int i0, int i1, int j0,
                                                    int j1, float* z, uint64_t ldz, int nrhs) {
    __shared__ float zs[RES_CHUNK][8];
    const int i = i0 + blockIdx.x * 256 + threadIdx.x;
    for (int e = threadIdx.x; e < (j1 - j0) * 8; e += 256) {
        const int j = e / 8, k = e % 8;
        zs[j][k] = k < nrhs ? z[(uint64_t)k * ldz + j0 + j] : 0.f;
    }
    __syncthreads();
    if (i >= i1) return;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = j0; j < j1; ++j) {
        const float av = a[(uint64_t)j * lda + i];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = fmaf(av, zs[j - j0][k], s[k]);
    }
    for (int k = 0; k < nrhs; ++k) z[(uint64_t)k * ldz + i] -= s[k];
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
    UH_LAUNCH(k_lu_panel, dim3(1), dim3(1024), 0, s, a, lda, n);
    UH_LAUNCH_CHECK("k_lu_panel");
    g_plain_launches++;
    return UH_OK;
}

static int trsm_common(const float* t, uint64_t ldt, float* x, uint64_t ldx, int n, int m, int kind, void* stream) {
    if (!t || !x || n <= 0 || m < 0) return fail(UH_ERR_ARG, "uh_trsm: bad arguments");
    if (n > TRSM_MAX_N) return fail(UH_ERR_ARG, "uh_trsm: n=%d above %d", n, TRSM_MAX_N);
    if (ldt < (uint64_t)n) return fail(UH_ERR_ARG, "uh_trsm: ldt too small");
    if (kind == 0 && ldx < (uint64_t)n) return fail(UH_ERR_ARG, "uh_trsm: ldx too small");
    if (kind == 1 && ldx < (uint64_t)m) return fail(UH_ERR_ARG, "uh_trsm: ldx too small");
    if (m == 0) return UH_OK;
    TrsmItem it{t, x, ldt, ldx, n, m, kind, 0};
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_TRSM, &it, sizeof(it), trsm_tiles(it));
        return UH_OK;
    }
    const size_t shm = (size_t)n * 33 * sizeof(float);
    UH_LAUNCH(k_trsm, dim3(trsm_tiles(it)), dim3(256), shm, (hipStream_t)stream, it);
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

extern "C" int uh_lu_solve_f32(const float* lu, uint64_t ld, int n, float* rhs, int nrhs, void* stream) {
    if (!lu || !rhs || n <= 0 || nrhs <= 0 || nrhs > 8 || ld < (uint64_t)n)
        return fail(UH_ERR_ARG, "uh_lu_solve_f32: bad arguments");
    int rc;
    hipStream_t s = plain_stream(stream, &rc);
    if (rc) return rc;
    const int B = RES_CHUNK;
    // forward: L z = r, block by block (diagonal solve, then update of the rows below)
    for (int b0 = 0; b0 < n; b0 += B) {
        const int bn = min(B, n - b0);
        rc = uh_trsm_lln(lu + (uint64_t)b0 * ld + b0, ld, rhs + b0, (uint64_t)n, bn, nrhs, s);
        if (rc) return rc;
        if (b0 + bn < n) {
            UH_LAUNCH(k_update_f32, dim3((n - b0 - bn + 255) / 256), dim3(256), 0, s, lu, ld, b0 + bn, n,
                               b0, b0 + bn, rhs, (uint64_t)n, nrhs);
            UH_LAUNCH_CHECK("k_update_f32");
            g_plain_launches++;
        }
    }
    // backward: U x = z, U unit upper, blocks from the bottom (diagonal solve on reversed
    // indices, then update of the rows above)
    for (int b1 = n; b1 > 0; b1 -= B) {
        const int b0 = max(0, b1 - B);
        const int bn = b1 - b0;
        TrsmItem it{lu + (uint64_t)b0 * ld + b0, rhs + b0, ld, (uint64_t)n, bn, nrhs, 2, 0};
        UH_LAUNCH(k_trsm, dim3(trsm_tiles(it)), dim3(256), (size_t)bn * 33 * sizeof(float), s, it);
        UH_LAUNCH_CHECK("k_trsm");
        g_plain_launches++;
        if (b0 > 0) {
            UH_LAUNCH(k_update_f32, dim3((b0 + 255) / 256), dim3(256), 0, s, lu, ld, 0, b0, b0, b1, rhs,
                               (uint64_t)n, nrhs);
            UH_LAUNCH_CHECK("k_update_f32");
            g_plain_launches++;
        }
    }
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
