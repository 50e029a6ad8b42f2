// imageFiltering subtask kernel for gfx950 (replaces imageFilter_cuda + prepareForLaunch,
// testSuite/imageFiltering/source/code/src/imageFilteringCUDA.cu:86-273).
//
// Semantics (both modes bit-exact to the reference CPU callback, imageFiltering.cpp:379-472):
//   out(i,j,c) = sum_{dy,dx} w[dy+r][dx+r] * in(clamp(i+dy), clamp(j+dx), 2-c)
// The input view is the subtask's halo'd region in inverted row order (reference
// imageFiltering.cpp:353-361); rows and columns are clamped to the image edges.
//
// Design. The reference stages every tile through a D2D copy into pitched texture memory and
// reads three separate texels per pixel. Here each workgroup computes a 256 x 16 output tile
// straight from the (HBM-resident) view:
//   1. the (16+2r) x (256+8) input window is read with dword loads (4 pixels = 12 bytes per
//      lane), de-interleaved with v_perm into three channel planes in LDS, and biased by
//      XOR 0x80 so that bytes can be used as signed int8;
//   2. each lane computes 4 consecutive pixels x 4 rows x 3 channels. For a filter row, the
//      2r+1 taps of pixel x+q lie in two aligned LDS dwords; v_dot4_i32_i8 against weight
//      vectors pre-shifted for each of the 4 pixel phases gives 2 instructions per
//      (pixel, channel, filter row) with no byte extraction;
//   3. the bias is undone by adding 128 * sum(w); mode 0 keeps the low byte (the reference's
//      char accumulators, i.e. arithmetic modulo 256), mode 1 shifts right arithmetically
//      and clamps to [0, 255].
// Radii 1 and 2 take this path; other radii use a direct per-pixel kernel.
#include "uh_internal.h"

#include <cstring>

namespace uh {

static constexpr int FT_W = 256;      // output tile width (pixels)
static constexpr int FT_H = 16;       // output tile height (rows)
static constexpr int FT_RPT = 4;      // output rows per lane
static constexpr int FT_THREADS = 256;
static constexpr int FT_LW = FT_W + 8;        // LDS plane width in pixels (4 either side)
static constexpr int FT_LWD = FT_LW / 4;      // ... in dwords

struct FilterView {
    int first_row, last_row, first_col, last_col;  // view extent in image coordinates
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) { return *(const uint32_t*)p; }

// Tile geometry shared by the plain and the persistent kernels.
__device__ __forceinline__ void filter_tile_origin(const uh_imagefilter_args& a, uint32_t tile,
                                                   int& x0, int& y0) {
    const int tiles_x = (a.cols + FT_W - 1) / FT_W;
    x0 = a.col0 + (int)(tile % (uint32_t)tiles_x) * FT_W;
    y0 = a.row0 + (int)(tile / (uint32_t)tiles_x) * FT_H;
}

// Stage the input window of one tile into three biased channel planes.
// plane[k][row][dword], row = image row (y0 - R + row) clamped to the view, LDS column c holds
// image column (x0 - 4 + c) clamped to the view.
template <int R>
__device__ __forceinline__ void filter_stage(const uh_imagefilter_args& a, const FilterView& v,
                                             int x0, int y0, uint32_t* lds) {
    constexpr int LROWS = FT_H + 2 * R;
    constexpr int GROUPS = FT_LW / 4;              // 4-pixel groups per LDS row
    const uint8_t* in = a.in;
    for (int g = threadIdx.x; g < LROWS * GROUPS; g += FT_THREADS) {
        const int lr = g / GROUPS;
        const int gc = g - lr * GROUPS;
        const int m = clampi(y0 - R + lr, v.first_row, v.last_row);
        const uint8_t* row = in + (uint64_t)(a.view_last_row - m) * a.in_stride;
        const int p = x0 - 4 + gc * 4;             // first image column of the group
        uint32_t p0, p1, p2;
        if (p >= v.first_col + 1 && p + 4 <= v.last_col) {
            // Fast path: 12 bytes of 4 in-view pixels; the aligned 16-byte window around them
            // stays inside the view row (pixels p-1 and p+4 are in the view).
            const uint8_t* src = row + (uint64_t)(p - v.first_col) * 3;
            const uintptr_t ad = (uintptr_t)src;
            const uint8_t* base = (const uint8_t*)(ad & ~(uintptr_t)3);
            const uint32_t sh = (uint32_t)(ad & 3u);
            const uint32_t d0 = ld_u32(base), d1 = ld_u32(base + 4), d2 = ld_u32(base + 8);
            if (sh == 0) {
                p0 = d0; p1 = d1; p2 = d2;
            } else {
                const uint32_t d3 = ld_u32(base + 12);
                p0 = __builtin_amdgcn_alignbyte(d1, d0, sh);
                p1 = __builtin_amdgcn_alignbyte(d2, d1, sh);
                p2 = __builtin_amdgcn_alignbyte(d3, d2, sh);
            }
        } else {
            uint32_t b[12];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n = clampi(p + q, v.first_col, v.last_col);
                const uint8_t* s = row + (uint64_t)(n - v.first_col) * 3;
                b[3 * q + 0] = s[0]; b[3 * q + 1] = s[1]; b[3 * q + 2] = s[2];
            }
            p0 = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
            p1 = b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24);
            p2 = b[8] | (b[9] << 8) | (b[10] << 16) | (b[11] << 24);
        }
        // De-interleave: plane k takes bytes k, k+3, k+6, k+9 of the 12-byte group.
        const uint32_t pl0 = (p0 & 0xffu) | (((p0 >> 24) & 0xffu) << 8) |
                             (((p1 >> 16) & 0xffu) << 16) | (((p2 >> 8) & 0xffu) << 24);
        const uint32_t pl1 = ((p0 >> 8) & 0xffu) | ((p1 & 0xffu) << 8) |
                             (((p1 >> 24) & 0xffu) << 16) | (((p2 >> 16) & 0xffu) << 24);
        const uint32_t pl2 = ((p0 >> 16) & 0xffu) | (((p1 >> 8) & 0xffu) << 8) |
                             ((p2 & 0xffu) << 16) | (((p2 >> 24) & 0xffu) << 24);
        const int idx = lr * FT_LWD + gc;
        lds[0 * LROWS * FT_LWD + idx] = pl0 ^ 0x80808080u;
        lds[1 * LROWS * FT_LWD + idx] = pl1 ^ 0x80808080u;
        lds[2 * LROWS * FT_LWD + idx] = pl2 ^ 0x80808080u;
    }
}

// Weight vectors: pixel x+q of lane l (x = x0 + 4l, LDS dword l+1 holds pixels x..x+3) needs
// LDS columns 4l+4+q-R .. 4l+4+q+R. Column c (relative to 4l) lives in dword c/4, byte c%4;
// for R <= 2 the span touches dwords d0 = (4+q-R)/4 and d0+1. wv[dy][q][0] / [1] pack the
// taps of filter row dy that fall into those two dwords (zero bytes elsewhere). They are packed
// once per work item on the host (pack_filter_weights) and read as uniform values per tile.
static void pack_filter_weights(const int8_t* w, int R, uint32_t* wv, int* wsum) {
    const int D = 2 * R + 1;
    int s = 0;
    for (int dy = 0; dy < D; ++dy) {
        for (int q = 0; q < 4; ++q) {
            const int c0 = 4 + q - R;          // first LDS column (relative to 4l)
            const int d0 = c0 / 4;             // first dword (relative to l)
            uint32_t w0 = 0, w1 = 0;
            for (int dx = 0; dx < D; ++dx) {
                const int c = c0 + dx;
                const uint32_t wb = (uint32_t)(uint8_t)w[dy * D + dx];
                if (c / 4 == d0) w0 |= wb << ((c % 4) * 8);
                else w1 |= wb << ((c % 4) * 8);
            }
            wv[(dy * 4 + q) * 2 + 0] = w0;
            wv[(dy * 4 + q) * 2 + 1] = w1;
        }
        for (int dx = 0; dx < D; ++dx) s += w[dy * D + dx];
    }
    *wsum = s;
}

template <int R>
__device__ __forceinline__ void filter_compute(const FilterItem& it, int x0, int y0,
                                               const uint32_t* lds) {
    constexpr int D = 2 * R + 1;
    constexpr int LROWS = FT_H + 2 * R;
    const uh_imagefilter_args& a = it.a;
    uint32_t wv[D][4][2];
#pragma unroll
    for (int dy = 0; dy < D; ++dy)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            wv[dy][q][0] = it.wv[(dy * 4 + q) * 2 + 0];
            wv[dy][q][1] = it.wv[(dy * 4 + q) * 2 + 1];
        }
    const int wsum = it.wsum;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ry0 = wave * FT_RPT;               // first output row of this lane (tile-relative)
    int acc[FT_RPT][4][3];
#pragma unroll
    for (int r = 0; r < FT_RPT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[r][q][k] = 0;

    // LDS rows ry0 .. ry0 + RPT - 1 + 2R feed output rows ry0 .. ry0 + RPT - 1.
#pragma unroll
    for (int lr = 0; lr < FT_RPT + 2 * R; ++lr) {
        uint32_t e[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t* pr = lds + k * LROWS * FT_LWD + (ry0 + lr) * FT_LWD + lane;
            e[k][0] = pr[0]; e[k][1] = pr[1]; e[k][2] = pr[2];
        }
#pragma unroll
        for (int r = 0; r < FT_RPT; ++r) {
            const int dy = lr - r;
            if (dy < 0 || dy >= D) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int d0 = (4 + q - R) / 4;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    int s = acc[r][q][k];
                    s = __builtin_amdgcn_sdot4((int)e[k][d0], (int)wv[dy][q][0], s, false);
                    if ((4 + q + R) / 4 != d0)   // span reaches the next dword
                        s = __builtin_amdgcn_sdot4((int)e[k][d0 + 1], (int)wv[dy][q][1], s, false);
                    acc[r][q][k] = s;
                }
            }
        }
    }

    // Epilogue: undo the XOR-0x80 bias, apply the mode, swap channels (out c <- in 2-c).
    const int bias = 128 * wsum;
    const int x = x0 + 4 * lane;
    const int cend = a.col0 + a.cols;
    const int rend = a.row0 + a.rows;
#pragma unroll
    for (int r = 0; r < FT_RPT; ++r) {
        const int y = y0 + ry0 + r;
        if (y >= rend || x >= cend) continue;
        uint32_t ob[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int v = acc[r][q][2 - c] + bias;
                uint32_t b;
                if (a.mode == 0) b = (uint32_t)v & 0xffu;
                else b = (uint32_t)clampi(v >> a.shift, 0, 255);
                ob[3 * q + c] = b;
            }
        }
        uint8_t* dst = a.out + (uint64_t)(y - a.row0) * a.out_stride + (uint64_t)(x - a.col0) * 3;
        if (x + 4 <= cend && (((uintptr_t)dst) & 3u) == 0) {
            uint32_t* d32 = (uint32_t*)dst;
            d32[0] = ob[0] | (ob[1] << 8) | (ob[2] << 16) | (ob[3] << 24);
            d32[1] = ob[4] | (ob[5] << 8) | (ob[6] << 16) | (ob[7] << 24);
            d32[2] = ob[8] | (ob[9] << 8) | (ob[10] << 16) | (ob[11] << 24);
        } else {
            const int np = (cend - x) < 4 ? (cend - x) : 4;
            for (int b = 0; b < 3 * np; ++b) dst[b] = (uint8_t)ob[b];
        }
    }
}

template <int R>
__device__ __forceinline__ void filter_tile(const FilterItem& it, uint32_t tile, uint32_t* lds) {
    const uh_imagefilter_args& a = it.a;
    int x0, y0;
    filter_tile_origin(a, tile, x0, y0);
    FilterView v;
    v.first_col = a.view_first_col;
    v.last_col = a.view_first_col + (int)a.pad - 1;      // pad carries the view width
    v.last_row = a.view_last_row;
    // view height: the view covers rows [max(row0-r,0), min(row0+rows-1+r, H-1)]
    v.first_row = a.row0 - a.radius < 0 ? 0 : a.row0 - a.radius;
    filter_stage<R>(a, v, x0, y0, lds);
    __syncthreads();
    filter_compute<R>(it, x0, y0, lds);
}

// Direct kernel for any radius <= 15: one output pixel per thread.
__device__ __forceinline__ void filter_pixel_direct(const uh_imagefilter_args& a, int i, int j) {
    const int r = a.radius, D = 2 * r + 1;
    const int first_row = a.row0 - r < 0 ? 0 : a.row0 - r;
    const int last_row = a.view_last_row;
    const int first_col = a.view_first_col;
    const int last_col = a.view_first_col + (int)a.pad - 1;
    int acc[3] = {0, 0, 0};
    for (int dy = -r; dy <= r; ++dy) {
        const int m = clampi(i + dy, first_row, last_row);
        const uint8_t* row = a.in + (uint64_t)(a.view_last_row - m) * a.in_stride;
        for (int dx = -r; dx <= r; ++dx) {
            const int n = clampi(j + dx, first_col, last_col);
            const uint8_t* px = row + (uint64_t)(n - first_col) * 3;
            const int w = a.weights[(dy + r) * D + (dx + r)];
            acc[0] += w * px[0]; acc[1] += w * px[1]; acc[2] += w * px[2];
        }
    }
    uint8_t* dst = a.out + (uint64_t)(i - a.row0) * a.out_stride + (uint64_t)(j - a.col0) * 3;
    for (int c = 0; c < 3; ++c) {
        const int v = acc[2 - c];
        dst[c] = a.mode == 0 ? (uint8_t)(v & 0xff) : (uint8_t)clampi(v >> a.shift, 0, 255);
    }
}

__device__ __forceinline__ void filter_tile_direct(const uh_imagefilter_args& a, uint32_t tile) {
    int x0, y0;
    filter_tile_origin(a, tile, x0, y0);
    for (int t = threadIdx.x; t < FT_W * FT_H; t += FT_THREADS) {
        const int i = y0 + t / FT_W, j = x0 + t % FT_W;
        if (i < a.row0 + a.rows && j < a.col0 + a.cols) filter_pixel_direct(a, i, j);
    }
}

__device__ __forceinline__ void filter_dispatch(const FilterItem& it, uint32_t tile, uint32_t* lds) {
    if (it.a.radius == 2) filter_tile<2>(it, tile, lds);
    else if (it.a.radius == 1) filter_tile<1>(it, tile, lds);
    else filter_tile_direct(it.a, tile);
}

static constexpr int FT_LDS_DWORDS = 3 * (FT_H + 2 * 2) * FT_LWD;

__global__ void __launch_bounds__(FT_THREADS) k_imagefilter(FilterItem it) {
    __shared__ uint32_t lds[FT_LDS_DWORDS];
    filter_dispatch(it, blockIdx.x, lds);
}

__global__ void __launch_bounds__(FT_THREADS) k_imagefilter_persistent(const uint8_t* queue) {
    __shared__ uint32_t lds[FT_LDS_DWORDS];
    __shared__ uint32_t s_tile;
    QueueHeader* h = (QueueHeader*)queue;
    const FilterItem* items = (const FilterItem*)(queue + h->items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + h->prefix_offset);
    const uint32_t total = h->total_tiles, n_items = h->n_items;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&h->next_tile, 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= total) break;
        uint32_t item, local;
        queue_locate(prefix, n_items, t, item, local);
        filter_dispatch(items[item], local, lds);
        __syncthreads();   // LDS and s_tile are reused by the next tile
    }
}

uint32_t filter_tiles(const FilterItem& it) {
    return (uint32_t)(((it.a.cols + FT_W - 1) / FT_W) * ((it.a.rows + FT_H - 1) / FT_H));
}

int filter_persistent(const uint8_t* queue, hipStream_t s) {
    UH_LAUNCH(k_imagefilter_persistent, dim3(cu_count() * 4), dim3(FT_THREADS), 0, s, queue);
    UH_LAUNCH_CHECK("k_imagefilter_persistent");
    g_persistent_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

// Validates the geometry a launch assumes. Returns the view width in pixels via *view_cols.
static int filter_validate(const uh_imagefilter_args* a, int* view_cols) {
    if (!a || !a->in || !a->out) return fail(UH_ERR_ARG, "uh_imagefilter: null pointer");
    if (a->radius < 0 || a->radius > 15) return fail(UH_ERR_ARG, "uh_imagefilter: radius %d", a->radius);
    if (a->mode != 0 && a->mode != 1) return fail(UH_ERR_ARG, "uh_imagefilter: mode %d", a->mode);
    if (a->shift < 0 || a->shift > 30) return fail(UH_ERR_ARG, "uh_imagefilter: shift %d", a->shift);
    if (a->height <= 0 || a->width <= 0 || a->rows <= 0 || a->cols <= 0)
        return fail(UH_ERR_ARG, "uh_imagefilter: empty geometry");
    if (a->row0 < 0 || a->col0 < 0 || a->row0 + a->rows > a->height || a->col0 + a->cols > a->width)
        return fail(UH_ERR_ARG, "uh_imagefilter: tile outside image");
    const int r = a->radius;
    const int lr = a->row0 + a->rows - 1 + r >= a->height ? a->height - 1 : a->row0 + a->rows - 1 + r;
    const int fc = a->col0 - r < 0 ? 0 : a->col0 - r;
    const int lc = a->col0 + a->cols - 1 + r >= a->width ? a->width - 1 : a->col0 + a->cols - 1 + r;
    if (a->view_last_row != lr || a->view_first_col != fc)
        return fail(UH_ERR_ARG, "uh_imagefilter: view origin (%d,%d) != halo (%d,%d)",
                    a->view_last_row, a->view_first_col, lr, fc);
    if ((uint64_t)(lc - fc + 1) * 3 > a->in_stride || (uint64_t)a->cols * 3 > a->out_stride)
        return fail(UH_ERR_ARG, "uh_imagefilter: stride smaller than a row");
    *view_cols = lc - fc + 1;
    // The weighted sum must fit int32 for mode 1 (always true for |w| <= 127, r <= 15).
    return UH_OK;
}

extern "C" int uh_imagefilter(const uh_imagefilter_args* args, void* stream) {
    int vcols = 0;
    int rc = filter_validate(args, &vcols);
    if (rc) return rc;
    FilterItem it;
    memset(&it, 0, sizeof(it));
    it.a = *args;
    it.a.pad = vcols;   // the kernels read the view width from pad
    if (it.a.radius <= 2) pack_filter_weights(it.a.weights, it.a.radius, it.wv, &it.wsum);
    const uint32_t tiles = filter_tiles(it);
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_FILTER, &it, sizeof(it), tiles);
        return UH_OK;
    }
    UH_LAUNCH(k_imagefilter, dim3(tiles), dim3(FT_THREADS), 0, (hipStream_t)stream, it);
    UH_LAUNCH_CHECK("k_imagefilter");
    g_plain_launches++;
    return UH_OK;
}
