// imageFiltering subtask kernel for gfx950 (replaces imageFilter_cuda + prepareForLaunch,
// testSuite/imageFiltering/source/code/src/imageFilteringCUDA.cu:86-273).
//
// Semantics (both modes bit-exact to the reference CPU callback, imageFiltering.cpp:379-472):
//   out(i,j,c) = sum_{dy,dx} w[dy+r][dx+r] * in(clamp(i+dy), clamp(j+dx), 2-c)
// The input view is the subtask's halo'd region in inverted row order (reference
// imageFiltering.cpp:353-361); rows and columns are clamped to the image edges.
//
// Design. The reference stages every tile through a D2D copy into pitched texture memory and
// reads three separate texels per pixel. Here each workgroup computes 256 x 16 output tiles
// straight from the (HBM-resident) view, several tiles per workgroup:
//   1. fetch: the (16+2r) x (256+8) input window is read into registers, 4 pixels (12 bytes)
//      per group, each group with one 16-byte global load from the dword-aligned address below
//      it; all loads of a tile are issued back to back;
//   2. store: v_alignbyte extracts the 12 bytes, v_perm de-interleaves them into three channel
//      planes, XOR 0x80 biases the bytes so that they can be used as signed int8, and the
//      planes go to LDS;
//   3. compute: each lane computes 4 consecutive pixels x 4 rows x 3 channels. For a filter row,
//      the 2r+1 taps of pixel x+q lie in two aligned LDS dwords; v_dot4_i32_i8 against weight
//      vectors pre-shifted for each of the 4 pixel phases gives 2 instructions per (pixel,
//      channel, filter row) with no byte extraction. The accumulators start at 128 * sum(w),
//      which undoes the bias; mode 0 keeps the low byte (the reference's char accumulators,
//      i.e. arithmetic modulo 256), mode 1 shifts right arithmetically and clamps to [0, 255].
// The per-item arguments and weight vectors are scalar loads (SGPRs); image loads and stores use
// global-space pointers, so LDS waits never wait for HBM traffic.
// Radii 1 and 2 take this path; other radii use a direct per-pixel path.
#include "uh_internal.h"

#include <cstddef>
#include <cstring>

namespace uh {

static constexpr int FT_W = 256;      // output tile width (pixels)
static constexpr int FT_H = 16;       // output tile height (rows)
static constexpr int FT_RPT = 4;      // output rows per lane
static constexpr int FT_THREADS = 256;
#ifndef FT_WAVES_PER_EU
#define FT_WAVES_PER_EU 4             // occupancy target: 4 waves per SIMD (<= 128 VGPRs)
#endif
#ifndef FT_PREFETCH
#define FT_PREFETCH 0                 // 1: fetch tile n+1 into registers while tile n is computed
#endif
static constexpr int FT_LW = FT_W + 8;        // LDS plane width in pixels (4 either side)
static constexpr int FT_LWD = FT_LW / 4;      // ... in dwords = 4-pixel groups per staged row

template <int R>
struct FtGeom {
    static constexpr int LROWS = FT_H + 2 * R;                 // staged input rows
    static constexpr int GROUPS = LROWS * FT_LWD;              // 4-pixel groups staged per tile
    static constexpr int ITERS = (GROUPS + FT_THREADS - 1) / FT_THREADS;
};
static constexpr int FT_MAX_ITERS = FtGeom<2>::ITERS;
static constexpr int FT_LDS_DWORDS = 3 * FtGeom<2>::LROWS * FT_LWD;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));

// uh_imagefilter_args without its weight table: the values the register-tiled path needs.
struct FilterArgs {
    const uint8_t* in;  uint64_t in_stride;
    uint8_t* out;       uint64_t out_stride;
    int32_t row0, col0, rows, cols;
    int32_t view_last_row, view_first_col;
    int32_t height, width;
    int32_t radius, mode, shift, pad;
};
static_assert(offsetof(uh_imagefilter_args, weights) == sizeof(FilterArgs), "FilterArgs is the head of the args");
static_assert(offsetof(uh_imagefilter_args, pad) == offsetof(FilterArgs, pad), "FilterArgs layout");

__device__ __forceinline__ FilterArgs head_of(const uh_imagefilter_args& a) {
    FilterArgs h;
    h.in = a.in; h.in_stride = a.in_stride; h.out = a.out; h.out_stride = a.out_stride;
    h.row0 = a.row0; h.col0 = a.col0; h.rows = a.rows; h.cols = a.cols;
    h.view_last_row = a.view_last_row; h.view_first_col = a.view_first_col;
    h.height = a.height; h.width = a.width;
    h.radius = a.radius; h.mode = a.mode; h.shift = a.shift; h.pad = a.pad;
    return h;
}

struct FilterView {
    int first_row, last_row, first_col, last_col;  // view extent in image coordinates
};

// One tile's input window in registers, between its global loads and its LDS stores. Group i of
// a lane holds the four dwords that start at the dword-aligned address at or below the group's
// first byte; bits 2i..2i+1 of `sh` give the group's byte offset within them.
struct FilterFetch {
    uint32_t d[FT_MAX_ITERS][4];
    uint32_t sh;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ FilterView filter_view(const FilterArgs& a) {
    FilterView v;
    v.first_col = a.view_first_col;
    v.last_col = a.view_first_col + (int)a.pad - 1;      // pad carries the view width
    v.last_row = a.view_last_row;
    // the view covers rows [max(row0 - r, 0), min(row0 + rows - 1 + r, H - 1)]
    v.first_row = a.row0 - a.radius < 0 ? 0 : a.row0 - a.radius;
    return v;
}

// Tile geometry shared by the plain and the persistent kernels.
__device__ __forceinline__ void filter_tile_origin(const FilterArgs& a, uint32_t tile, int& x0, int& y0) {
    const int tiles_x = (a.cols + FT_W - 1) / FT_W;
    x0 = a.col0 + (int)(tile % (uint32_t)tiles_x) * FT_W;
    y0 = a.row0 + (int)(tile / (uint32_t)tiles_x) * FT_H;
}

// Fetch the input window of one tile. Group g (= staged row lr, group column gc) covers image
// columns x0 - 4 + 4gc .. +3 of image row y0 - R + lr, both clamped to the view.
template <int R>
__device__ __forceinline__ void filter_fetch(const FilterArgs& a, uint32_t tile, FilterFetch& f) {
    using G = FtGeom<R>;
    int x0, y0;
    filter_tile_origin(a, tile, x0, y0);
    const FilterView v = filter_view(a);
    const UH_GLOBAL uint8_t* in = as_global(a.in);
    f.sh = 0;
#pragma unroll
    for (int i = 0; i < G::ITERS; ++i) {
        const int g = threadIdx.x + i * FT_THREADS;
        if (i == G::ITERS - 1 && g >= G::GROUPS) break;
        const int lr = g / FT_LWD;
        const int gc = g - lr * FT_LWD;
        const int m = clampi(y0 - R + lr, v.first_row, v.last_row);
        const UH_GLOBAL uint8_t* row = in + (uint64_t)(a.view_last_row - m) * a.in_stride;
        const int p = x0 - 4 + gc * 4;             // first image column of the group
        if (p - 1 >= v.first_col && p + 5 <= v.last_col) {
            // Pixels p-1 .. p+5 are in the view row, so the 16 bytes from the dword-aligned
            // address at or below pixel p (which start at most 3 bytes before it) stay inside.
            const UH_GLOBAL uint8_t* src = row + (uint64_t)(p - v.first_col) * 3;
            const uint32_t sh = (uint32_t)((uintptr_t)src & 3u);
            const UH_GLOBAL uint32_t* base = (const UH_GLOBAL uint32_t*)(src - sh);
            f.d[i][0] = base[0]; f.d[i][1] = base[1]; f.d[i][2] = base[2]; f.d[i][3] = base[3];
            f.sh |= sh << (2 * i);
        } else {
            // Group at an image or view edge: clamped per-pixel byte loads, offset 0.
            uint32_t b[12];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int n = clampi(p + k, v.first_col, v.last_col);
                const UH_GLOBAL uint8_t* s = row + (uint64_t)(n - v.first_col) * 3;
                b[3 * k + 0] = s[0]; b[3 * k + 1] = s[1]; b[3 * k + 2] = s[2];
            }
            f.d[i][0] = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
            f.d[i][1] = b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24);
            f.d[i][2] = b[8] | (b[9] << 8) | (b[10] << 16) | (b[11] << 24);
            f.d[i][3] = 0;
        }
    }
}

// Store a fetched window as three biased channel planes: plane[k][lr][gc] holds channel k of
// the group's 4 pixels (LDS dword index lr * FT_LWD + gc == g).
template <int R>
__device__ __forceinline__ void filter_store(const FilterFetch& f, uint32_t* lds) {
    using G = FtGeom<R>;
#pragma unroll
    for (int i = 0; i < G::ITERS; ++i) {
        const int g = threadIdx.x + i * FT_THREADS;
        if (i == G::ITERS - 1 && g >= G::GROUPS) break;
        const uint32_t s = (f.sh >> (2 * i)) & 3u;
        const uint32_t p0 = __builtin_amdgcn_alignbyte(f.d[i][1], f.d[i][0], s);
        const uint32_t p1 = __builtin_amdgcn_alignbyte(f.d[i][2], f.d[i][1], s);
        const uint32_t p2 = __builtin_amdgcn_alignbyte(f.d[i][3], f.d[i][2], s);
        // v_perm selects bytes of {S0:S1} (S1 = bytes 0-3, S0 = bytes 4-7; 0x0c gives zero).
        // Plane k takes bytes k, k+3, k+6, k+9 of the 12-byte group p0:p1:p2.
        const uint32_t t0 = __builtin_amdgcn_perm(p1, p0, 0x0c060300u);   // p0.0 p0.3 p1.2
        const uint32_t t1 = __builtin_amdgcn_perm(p1, p0, 0x0c070401u);   // p0.1 p1.0 p1.3
        const uint32_t t2 = __builtin_amdgcn_perm(p1, p0, 0x0c0c0502u);   // p0.2 p1.1
        const uint32_t pl0 = __builtin_amdgcn_perm(p2, t0, 0x05020100u);  // + p2.1
        const uint32_t pl1 = __builtin_amdgcn_perm(p2, t1, 0x06020100u);  // + p2.2
        const uint32_t pl2 = __builtin_amdgcn_perm(p2, t2, 0x07040100u);  // + p2.0 p2.3
        lds[0 * G::LROWS * FT_LWD + g] = pl0 ^ 0x80808080u;
        lds[1 * G::LROWS * FT_LWD + g] = pl1 ^ 0x80808080u;
        lds[2 * G::LROWS * FT_LWD + g] = pl2 ^ 0x80808080u;
    }
}

// Weight vectors: pixel x+q of lane l (x = x0 + 4l, LDS dword l+1 holds pixels x..x+3) needs
// LDS columns 4l+4+q-R .. 4l+4+q+R. Column c (relative to 4l) lives in dword c/4, byte c%4;
// for R <= 2 the span touches dwords d0 = (4+q-R)/4 and d0+1. wv[dy][q][0] / [1] pack the
// taps of filter row dy that fall into those two dwords (zero bytes elsewhere). They are packed
// once per work item on the host and read as uniform values per tile.
static void pack_filter_weights(const int8_t* w, int R, FilterWeights* out) {
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
            out->wv[(dy * 4 + q) * 2 + 0] = w0;
            out->wv[(dy * 4 + q) * 2 + 1] = w1;
        }
        for (int dx = 0; dx < D; ++dx) s += w[dy * D + dx];
    }
    out->wsum = s;
}

// Low bytes of four values packed into one dword (byte 0 = b0).
__device__ __forceinline__ uint32_t pack4(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) {
    return __builtin_amdgcn_perm(b1, b0, 0x0c0c0400u) | __builtin_amdgcn_perm(b3, b2, 0x04000c0cu);
}

template <int R, int MODE>
__device__ __forceinline__ void filter_compute(const FilterArgs& a, const FilterWeights& w, uint32_t tile,
                                               const uint32_t* lds) {
    constexpr int D = 2 * R + 1;
    constexpr int LROWS = FtGeom<R>::LROWS;
    int x0, y0;
    filter_tile_origin(a, tile, x0, y0);
    const int bias = 128 * w.wsum;               // undoes the XOR-0x80 bias of the inputs

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ry0 = wave * FT_RPT;               // first output row of this lane (tile-relative)
    int acc[FT_RPT][4][3];
#pragma unroll
    for (int r = 0; r < FT_RPT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[r][q][k] = bias;

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
                const int wi = (dy * 4 + q) * 2;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    int s = acc[r][q][k];
                    s = __builtin_amdgcn_sdot4((int)e[k][d0], (int)w.wv[wi], s, false);
                    if ((4 + q + R) / 4 != d0)   // span reaches the next dword
                        s = __builtin_amdgcn_sdot4((int)e[k][d0 + 1], (int)w.wv[wi + 1], s, false);
                    acc[r][q][k] = s;
                }
            }
        }
    }

    // Epilogue: apply the mode, swap channels (out c <- in 2-c), store 12 bytes per row.
    const int shift = a.shift;
    const int x = x0 + 4 * lane;
    const int cend = a.col0 + a.cols;
    const int rend = a.row0 + a.rows;
    UH_GLOBAL uint8_t* out = as_global(a.out);
#pragma unroll
    for (int r = 0; r < FT_RPT; ++r) {
        const int y = y0 + ry0 + r;
        if (y >= rend || x >= cend) continue;
        uint32_t ob[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int v = acc[r][q][2 - c];
                ob[3 * q + c] = MODE == 0 ? (uint32_t)v : (uint32_t)clampi(v >> shift, 0, 255);
            }
        }
        UH_GLOBAL uint8_t* dst = out + (uint64_t)(y - a.row0) * a.out_stride + (uint64_t)(x - a.col0) * 3;
        if (x + 4 <= cend && (((uintptr_t)dst) & 3u) == 0) {
            UH_GLOBAL uint32_t* d32 = (UH_GLOBAL uint32_t*)dst;
            d32[0] = pack4(ob[0], ob[1], ob[2], ob[3]);
            d32[1] = pack4(ob[4], ob[5], ob[6], ob[7]);
            d32[2] = pack4(ob[8], ob[9], ob[10], ob[11]);
        } else {
            const int nb = 3 * ((cend - x) < 4 ? (cend - x) : 4);
#pragma unroll
            for (int b = 0; b < 12; ++b)
                if (b < nb) dst[b] = (uint8_t)ob[b];
        }
    }
}

// Direct path for any radius <= 15: one output pixel per thread.
__device__ __forceinline__ void filter_pixel_direct(const FilterArgs& a, const int8_t* weights, int i, int j) {
    const int r = a.radius, D = 2 * r + 1;
    const int first_row = a.row0 - r < 0 ? 0 : a.row0 - r;
    const int last_row = a.view_last_row;
    const int first_col = a.view_first_col;
    const int last_col = a.view_first_col + (int)a.pad - 1;
    const UH_GLOBAL uint8_t* in = as_global(a.in);
    int acc[3] = {0, 0, 0};
    for (int dy = -r; dy <= r; ++dy) {
        const int m = clampi(i + dy, first_row, last_row);
        const UH_GLOBAL uint8_t* row = in + (uint64_t)(a.view_last_row - m) * a.in_stride;
        for (int dx = -r; dx <= r; ++dx) {
            const int n = clampi(j + dx, first_col, last_col);
            const UH_GLOBAL uint8_t* px = row + (uint64_t)(n - first_col) * 3;
            const int w = weights[(dy + r) * D + (dx + r)];
            acc[0] += w * px[0]; acc[1] += w * px[1]; acc[2] += w * px[2];
        }
    }
    UH_GLOBAL uint8_t* dst = as_global(a.out) + (uint64_t)(i - a.row0) * a.out_stride + (uint64_t)(j - a.col0) * 3;
    for (int c = 0; c < 3; ++c) {
        const int v = acc[2 - c];
        dst[c] = a.mode == 0 ? (uint8_t)(v & 0xff) : (uint8_t)clampi(v >> a.shift, 0, 255);
    }
}

__device__ __forceinline__ void filter_tile_direct(const FilterArgs& a, const int8_t* weights, uint32_t tile) {
    int x0, y0;
    filter_tile_origin(a, tile, x0, y0);
    for (int t = threadIdx.x; t < FT_W * FT_H; t += FT_THREADS) {
        const int i = y0 + t / FT_W, j = x0 + t % FT_W;
        if (i < a.row0 + a.rows && j < a.col0 + a.cols) filter_pixel_direct(a, weights, i, j);
    }
}

// Kernel variants. Variant 0 dispatches on each item's radius and mode (uniform per tile; the
// direct path fetches and stages nothing); variants 1..4 are specialised for one radius and mode
// (V = 1 + 2 (R - 1) + MODE) and are launched only on items of that kind, which keeps the
// register allocation of the common case free of the other paths.
__host__ __device__ constexpr int filter_variant(int radius, int mode) {
    return (radius == 1 || radius == 2) ? 1 + 2 * (radius - 1) + mode : 0;
}

template <int V>
__device__ __forceinline__ void filter_fetch_v(const FilterArgs& a, uint32_t tile, FilterFetch& f) {
    if constexpr (V == 0) {
        if (a.radius == 2) filter_fetch<2>(a, tile, f);
        else if (a.radius == 1) filter_fetch<1>(a, tile, f);
    } else {
        filter_fetch<(V + 1) / 2>(a, tile, f);
    }
}

template <int V>
__device__ __forceinline__ void filter_store_v(const FilterArgs& a, const FilterFetch& f, uint32_t* lds) {
    if constexpr (V == 0) {
        if (a.radius == 2) filter_store<2>(f, lds);
        else if (a.radius == 1) filter_store<1>(f, lds);
    } else {
        filter_store<(V + 1) / 2>(f, lds);
    }
}

// weights: the item's full weight table (used by the direct path only)
template <int V>
__device__ __forceinline__ void filter_compute_v(const FilterArgs& a, const FilterWeights& w,
                                                 const int8_t* weights, uint32_t tile, const uint32_t* lds) {
    if constexpr (V == 0) {
        if (a.radius == 2) {
            if (a.mode == 0) filter_compute<2, 0>(a, w, tile, lds);
            else filter_compute<2, 1>(a, w, tile, lds);
        } else if (a.radius == 1) {
            if (a.mode == 0) filter_compute<1, 0>(a, w, tile, lds);
            else filter_compute<1, 1>(a, w, tile, lds);
        } else {
            filter_tile_direct(a, weights, tile);
        }
    } else {
        filter_compute<(V + 1) / 2, (V + 1) % 2>(a, w, tile, lds);
    }
}

// Plain launch: workgroup b computes tiles b, b + grid, b + 2 grid, ... of one item.
template <int V>
__global__ void __launch_bounds__(FT_THREADS) __attribute__((amdgpu_waves_per_eu(FT_WAVES_PER_EU)))
k_imagefilter(FilterItem it, uint32_t tiles) {
    __shared__ uint32_t lds[FT_LDS_DWORDS];
    const FilterArgs a = head_of(it.a);
#if FT_PREFETCH
    // tile n+1 is fetched into registers while tile n is computed
    uint32_t t = blockIdx.x;
    if (t >= tiles) return;
    FilterFetch f;
    filter_fetch_v<V>(a, t, f);
    for (;;) {
        filter_store_v<V>(a, f, lds);
        __syncthreads();
        const uint32_t tn = t + gridDim.x;
        if (tn < tiles) filter_fetch_v<V>(a, tn, f);
        filter_compute_v<V>(a, it.w, it.a.weights, t, lds);
        if (tn >= tiles) break;
        __syncthreads();          // LDS is restaged for the next tile
        t = tn;
    }
#else
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        FilterFetch f;
        filter_fetch_v<V>(a, t, f);
        filter_store_v<V>(a, f, lds);
        __syncthreads();
        filter_compute_v<V>(a, it.w, it.a.weights, t, lds);
        __syncthreads();          // LDS is restaged for the next tile
    }
#endif
}

// Persistent launch over a device work queue. Tiles are claimed from an atomic counter one claim
// ahead (the claim for the tile after next is issued before a tile is computed and published to
// the workgroup after it), and the next tile is looked up (item search over the LDS copy of the
// tile prefix, scalar loads of its arguments) while the current one is computed. Every workgroup
// leaves once a claim returns an index past the total, so the grid always drains.
static constexpr uint32_t FT_PREFIX_LDS = 1024;   // items whose tile prefix is searched in LDS

template <int V>
__global__ void __launch_bounds__(FT_THREADS) __attribute__((amdgpu_waves_per_eu(FT_WAVES_PER_EU)))
k_imagefilter_persistent(const uint8_t* queue) {
    __shared__ uint32_t lds[FT_LDS_DWORDS];
    __shared__ uint32_t s_tile[2];
    __shared__ uint32_t s_prefix[FT_PREFIX_LDS + 1];
    QueueHeader* h = (QueueHeader*)queue;
    const QueueHeader hd = uniform_load(h);       // n_items, total_tiles, offsets (not next_tile)
    const FilterItem* items = (const FilterItem*)(queue + hd.items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + hd.prefix_offset);
    const uint32_t total = hd.total_tiles, n_items = hd.n_items;
    const bool lds_prefix = stage_prefix(prefix, n_items, s_prefix, FT_PREFIX_LDS + 1);
    if (threadIdx.x == 0) {
        s_tile[0] = atomicAdd(&h->next_tile, 1u);
        s_tile[1] = atomicAdd(&h->next_tile, 1u);
    }
    __syncthreads();
    auto locate = [&](uint32_t tile, uint32_t& item, uint32_t& local) {
        if (lds_prefix) queue_locate_lds(s_prefix, n_items, tile, item, local);
        else queue_locate(prefix, n_items, tile, item, local);
    };
    uint32_t t = __builtin_amdgcn_readfirstlane(s_tile[0]);
    uint32_t tn = __builtin_amdgcn_readfirstlane(s_tile[1]);
    if (t >= total) return;
    uint32_t item, local;
    locate(t, item, local);
    FilterArgs a = uniform_load((const FilterArgs*)&items[item].a);
#if FT_PREFETCH
    // the input of tile n+1 is fetched into registers while tile n is computed
    FilterFetch f;
    filter_fetch_v<V>(a, local, f);
    for (;;) {
        filter_store_v<V>(a, f, lds);
        __syncthreads();
        uint32_t item_n = 0, local_n = 0;
        FilterArgs an = a;
        if (tn < total) {
            locate(tn, item_n, local_n);
            an = uniform_load((const FilterArgs*)&items[item_n].a);
            filter_fetch_v<V>(an, local_n, f);
        }
        uint32_t claim = 0;
        if (threadIdx.x == 0 && tn < total) claim = atomicAdd(&h->next_tile, 1u);
        {
            const FilterWeights w = uniform_load(&items[item].w);
            filter_compute_v<V>(a, w, items[item].a.weights, local, lds);
        }
        if (tn >= total) break;
        if (threadIdx.x == 0) s_tile[0] = claim;
        __syncthreads();          // LDS is restaged and s_tile is read below
        t = tn;
        tn = __builtin_amdgcn_readfirstlane(s_tile[0]);
        item = item_n;
        local = local_n;
        a = an;
    }
#else
    for (;;) {
        {
            FilterFetch f;
            filter_fetch_v<V>(a, local, f);
            filter_store_v<V>(a, f, lds);
        }
        uint32_t claim = total;
        if (threadIdx.x == 0 && tn < total) claim = atomicAdd(&h->next_tile, 1u);
        __syncthreads();
        // look up the next tile now; its scalar loads complete while this tile is computed
        uint32_t item_n = 0, local_n = 0;
        FilterArgs an = a;
        if (tn < total) {
            locate(tn, item_n, local_n);
            an = uniform_load((const FilterArgs*)&items[item_n].a);
        }
        {
            const FilterWeights w = uniform_load(&items[item].w);
            filter_compute_v<V>(a, w, items[item].a.weights, local, lds);
        }
        if (tn >= total) break;
        if (threadIdx.x == 0) s_tile[0] = claim;
        __syncthreads();          // LDS is restaged and s_tile is read below
        t = tn;
        tn = __builtin_amdgcn_readfirstlane(s_tile[0]);
        item = item_n;
        local = local_n;
        a = an;
    }
#endif
}

uint32_t filter_tiles(const FilterItem& it) {
    return (uint32_t)(((it.a.cols + FT_W - 1) / FT_W) * ((it.a.rows + FT_H - 1) / FT_H));
}

int filter_variant_for(uint32_t kinds) {
    // one specialised variant present and nothing else
    return (kinds && (kinds & (kinds - 1)) == 0 && !(kinds & 1u)) ? __builtin_ctz(kinds) : 0;
}

typedef void (*FilterPersistentFn)(const uint8_t*);
static const FilterPersistentFn kFilterPersistent[5] = {
    k_imagefilter_persistent<0>, k_imagefilter_persistent<1>, k_imagefilter_persistent<2>,
    k_imagefilter_persistent<3>, k_imagefilter_persistent<4>};
typedef void (*FilterPlainFn)(FilterItem, uint32_t);
static const FilterPlainFn kFilterPlain[5] = {
    k_imagefilter<0>, k_imagefilter<1>, k_imagefilter<2>, k_imagefilter<3>, k_imagefilter<4>};

// Workgroups of a persistent filter kernel resident per CU (occupancy queries, cached).
static int g_filter_wg[5] = {};

int filter_wg_per_cu() {
    if (!g_filter_wg[0]) {
        for (int v = 0; v < 5; ++v) {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kFilterPersistent[v], FT_THREADS, 0) != hipSuccess || b < 1)
                b = 2;
            g_filter_wg[v] = b;
        }
    }
    return g_filter_wg[0];
}

int filter_persistent(const uint8_t* queue, hipStream_t s, uint32_t kinds) {
    const int v = filter_variant_for(kinds);
    filter_wg_per_cu();
    UH_LAUNCH(kFilterPersistent[v], dim3(cu_count() * g_filter_wg[v]), dim3(FT_THREADS), 0, s, queue);
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
    if (it.a.radius == 1 || it.a.radius == 2) pack_filter_weights(it.a.weights, it.a.radius, &it.w);
    const uint32_t tiles = filter_tiles(it);
    const int v = filter_variant(it.a.radius, it.a.mode);
    if (Batch* b = as_batch(stream)) {
        record_item(b, FAM_FILTER, &it, sizeof(it), tiles, 1u << v);
        return UH_OK;
    }
    // two tiles per workgroup
    const uint32_t grid = (tiles + 1) / 2;
    UH_LAUNCH(kFilterPlain[v], dim3(grid), dim3(FT_THREADS), 0, (hipStream_t)stream, it, tiles);
    UH_LAUNCH_CHECK("k_imagefilter");
    g_plain_launches++;
    return UH_OK;
}
