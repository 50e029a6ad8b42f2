// imageFiltering subtask kernel for gfx950 (replaces imageFilter_cuda + prepareForLaunch,
// testSuite/imageFiltering/source/code/src/imageFilteringCUDA.cu:86-273).
//
// Semantics (both modes bit-exact to the reference CPU callback, imageFiltering.cpp:379-472):
//   out(i,j,c) = sum_{dy,dx} w[dy+r][dx+r] * in(clamp(i+dy), clamp(j+dx), 2-c)
// The input view is the subtask's halo'd region in inverted row order (reference
// imageFiltering.cpp:353-361); rows and columns are clamped to the image edges.
//
// Design. The reference stages every tile through a D2D copy into pitched texture memory and
// reads three separate texels per pixel. Here the tiles are computed straight from the
// (HBM-resident) view by one of two engines. The default int8 MFMA engine (the fm_* functions
// below) computes 128 x 32 tiles as banded products over the raw interleaved bytes, through a
// three-buffer direct-to-LDS window ring and an LDS output staging area. The v_dot4 engine
// computes 256 x 16 tiles, several per workgroup, as a pipeline:
//   1. stage: the (16+2r)-row input window (264 pixels from x0 - 4) moves from HBM to LDS with
//      16-byte direct-to-LDS buffer loads, one tile ahead of the compute, holding no registers;
//   2. de-interleave: v_alignbyte extracts each 4-pixel group's 12 bytes, v_perm splits them into
//      three channel planes, XOR 0x80 biases the bytes so that they can be used as signed int8;
//   3. compute: each lane computes 4 consecutive pixels x 4 rows x 3 channels. For a filter row,
//      the 2r+1 taps of pixel x+q lie in two aligned LDS dwords; v_dot4_i32_i8 against weight
//      vectors pre-shifted for each of the 4 pixel phases gives 2 instructions per (pixel,
//      channel, filter row) with no byte extraction. The accumulators start at 128 * sum(w),
//      which undoes the bias; mode 0 keeps the low byte (the reference's char accumulators,
//      i.e. arithmetic modulo 256), mode 1 shifts right arithmetically and clamps to [0, 255].
// The per-item arguments and weight vectors are scalar loads (SGPRs); image stores use
// global-space pointers, so LDS waits never wait for HBM traffic.
// Radii 1 and 2 take this path; other radii use a direct per-pixel path.
#include <type_traits>
#include "uh_internal.h"

#include <cstddef>
#include <cstdlib>
#include <cstring>

namespace uh {

static constexpr int FT_W = 256;      // output tile width (pixels)
static constexpr int FT_H = 16;       // output tile height (rows)
static constexpr int FT_RPT = 4;      // output rows per lane
static constexpr int FT_THREADS = 256;
#ifndef FT_WAVES_PER_EU
#define FT_WAVES_PER_EU 4             // occupancy target: 4 waves per SIMD (<= 128 VGPRs)
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
// Tiles are numbered down each 256-pixel column strip, so consecutive tiles are vertical
// neighbours and share halo rows.
__device__ __forceinline__ void filter_tile_origin(const FilterArgs& a, uint32_t tile, int& x0, int& y0) {
    const uint32_t tiles_y = (uint32_t)((a.rows + FT_H - 1) / FT_H);
    x0 = a.col0 + (int)(tile / tiles_y) * FT_W;
    y0 = a.row0 + (int)(tile % tiles_y) * FT_H;
}

// Staged raw input. Row lr of a tile's window is the 816 bytes (51 16-byte chunks) from the
// 16-byte-aligned address at or below the window's first byte, pixel x0 - 4 of image row
// y0 - R + lr (rows clamped to the view). The chunks go straight from HBM to LDS
// (buffer_load_dwordx4 ... lds: no registers hold data in flight) through a buffer resource
// spanning the view, so chunks before or past the view read as zero instead of touching memory;
// the de-interleave below reads only pixels inside the view (columns clamped to it), which lie
// in the staged bytes.
static constexpr int FT_RAW_ROW = 816;
static constexpr int FT_RAW_CHUNKS = FT_RAW_ROW / 16;
// whole 1 KB wave destinations (64 lanes x 16 bytes), so a wave's last chunk group stays inside
static constexpr int FT_RAW_DWORDS = (FtGeom<2>::LROWS * FT_RAW_ROW + 1023) / 1024 * 256;

struct FilterLds {
    uint32_t planes[FT_LDS_DWORDS];
    uint32_t raw[FT_RAW_DWORDS];
};

// Offset from the view pointer of the first window byte of staged row lr (may be negative). The
// host keeps the view below 2 GiB and the stride below 2^24, so one full-rate 24-bit multiply
// gives the row offset.
template <int R>
__device__ __forceinline__ int filter_row_start(const FilterArgs& a, const FilterView& v, int x0, int y0, int lr) {
    const int m = clampi(y0 - R + lr, v.first_row, v.last_row);
    return (int)__umul24((uint32_t)(a.view_last_row - m), (uint32_t)a.in_stride) + (x0 - 4 - v.first_col) * 3;
}

// a * b for a, b < 2^24: the two full-rate 24-bit multiplies (low and high half of the product).
__device__ __forceinline__ uint64_t mul24_wide(uint32_t a, uint32_t b) {
    return (uint64_t)(a & 0xFFFFFFu) * (uint64_t)(b & 0xFFFFFFu);
}

template <int R>
__device__ __forceinline__ void filter_dma(const FilterArgs& a, int x0, int y0, uint32_t* raw) {
    constexpr int CHUNKS = FtGeom<R>::LROWS * FT_RAW_CHUNKS;
    const FilterView v = filter_view(a);
    // The view spans [in, in + extent) (rows in inverted order; the host keeps it below 2 GiB).
    // The resource starts at the 16-byte boundary at or below `in` and ends at the one at or
    // above the view's end: every 16-byte chunk holding a view byte is read whole (an aligned
    // chunk never leaves the page of its bytes), chunks past the end read as zero, and a chunk
    // wholly before the view (a window starting left of it) is redirected to offset 0.
    const int mis = (int)((uintptr_t)a.in & 15u);
    const int extent = (int)__umul24((uint32_t)(a.view_last_row - v.first_row), (uint32_t)a.in_stride) + a.pad * 3;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.in - mis), 0, (mis + extent + 15) & ~15, 0x00020000);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < (CHUNKS + FT_THREADS - 1) / FT_THREADS; ++i) {
        const int c0 = i * FT_THREADS + wave * 64;   // this wave's first chunk: LDS destination base
        if (c0 >= CHUNKS) break;
        const int c = c0 + lane;
        if (c < CHUNKS) {
            const int lr = c / FT_RAW_CHUNKS, ch = c - lr * FT_RAW_CHUNKS;
            const int st = filter_row_start<R>(a, v, x0, y0, lr) + mis;   // from the aligned base
            const int off = (st & ~15) + ch * 16;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rs, (__attribute__((address_space(3))) void*)((uint8_t*)raw + c0 * 16), 16,
                (uint32_t)(off < 0 ? 0 : off), 0, 0, 0);
        }
    }
}

// De-interleave the staged window into three biased channel planes: plane[k][lr][gc] holds
// channel k of the 4 pixels x0 - 4 + 4 gc .. +3 of staged row lr (LDS dword lr * FT_LWD + gc == g).
template <int R>
__device__ __forceinline__ void filter_unpack(const FilterArgs& a, int x0, int y0, const uint32_t* raw, uint32_t* lds) {
    using G = FtGeom<R>;
    const FilterView v = filter_view(a);
    const int mis = (int)((uintptr_t)a.in & 15u);
    const uint8_t* rawb = (const uint8_t*)raw;
#pragma unroll 2
    for (int i = 0; i < G::ITERS; ++i) {
        const int g = threadIdx.x + i * FT_THREADS;
        if (i == G::ITERS - 1 && g >= G::GROUPS) break;
        const int lr = g / FT_LWD;
        const int gc = g - lr * FT_LWD;
        const int st = filter_row_start<R>(a, v, x0, y0, lr);
        const int rb = lr * FT_RAW_ROW + ((st + mis) & 15);   // raw byte of pixel x0 - 4
        const int p = x0 - 4 + gc * 4;                            // first image column of the group
        uint32_t p0, p1, p2;
        if (p >= v.first_col && p + 3 <= v.last_col) {
            // 12 bytes at rb + 12 gc, inside the staged row: the four dwords from the one below
            const int o = rb + 12 * gc;
            const uint32_t* q = raw + (o >> 2);
            const uint32_t s = (uint32_t)o & 3u;
            const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3];
            p0 = __builtin_amdgcn_alignbyte(d1, d0, s);
            p1 = __builtin_amdgcn_alignbyte(d2, d1, s);
            p2 = __builtin_amdgcn_alignbyte(d3, d2, s);
        } else {
            // group at a view edge: pixels clamped to the view, read byte by byte
            uint32_t b[12];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int n = clampi(p + k, v.first_col, v.last_col);
                const uint8_t* px = rawb + rb + (n - (x0 - 4)) * 3;
                b[3 * k + 0] = px[0]; b[3 * k + 1] = px[1]; b[3 * k + 2] = px[2];
            }
            p0 = b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24);
            p1 = b[4] | (b[5] << 8) | (b[6] << 16) | (b[7] << 24);
            p2 = b[8] | (b[9] << 8) | (b[10] << 16) | (b[11] << 24);
        }
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

typedef int FilterAcc[FT_RPT][4][3];

// Dot products of one tile: lane l of wave w, pixels x0 + 4 l .. +3 of rows y0 + 4 w .. +3.
template <int R>
__device__ __forceinline__ void filter_accumulate(const FilterWeights& w, const uint32_t* lds, FilterAcc& acc) {
    constexpr int D = 2 * R + 1;
    constexpr int LROWS = FtGeom<R>::LROWS;
    const int bias = 128 * w.wsum;               // undoes the XOR-0x80 bias of the inputs
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ry0 = wave * FT_RPT;               // first output row of this lane (tile-relative)
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
}

// Epilogue: apply the mode, swap channels (out c <- in 2-c), store 12 bytes per row.
template <int MODE>
__device__ __forceinline__ void filter_epilogue(const FilterArgs& a, int x0, int y0, const FilterAcc& acc) {
    const int lane = threadIdx.x & 63;
    const int ry0 = (threadIdx.x >> 6) * FT_RPT;
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
        UH_GLOBAL uint8_t* dst = out + mul24_wide((uint32_t)(y - a.row0), (uint32_t)a.out_stride) + (uint32_t)(x - a.col0) * 3u;
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


__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }   // vmcnt(0)

// ---------------------------------------------------------------- MFMA engine (radius 1 and 2)
//
// For one filter row dy, the output bytes of an image row are a banded matrix product over the
// raw interleaved input bytes: output byte j (pixel x = j / 3, channel c = j % 3) takes input
// byte 3 (x + dx) + 2 - c = j + 3 dx + 2 - 2c for dx in [-R, R] (the B<->R swap included), so
//   out[y][j] = sum_dy sum_k in[y + dy][k] T_dy[k][j],   T_dy[j + 3dx + 2 - 2c(j)][j] = w[dy][dx].
// v_mfma_i32_32x32x32_i8 computes it for 32 output bytes (M) x 32 image rows (N): A = the band of
// T (32 outputs x 32 input bytes), B = 32 staged input bytes of 32 rows, read straight from the
// raw window in LDS (no de-interleave). A block of 32 outputs needs the 48 input bytes from 8
// before it, covered by two K steps of 32; the three blocks of a wave's 96 bytes (32 pixels) share
// their middle steps, so a filter row costs 4 B reads and 6 MFMAs per 32 rows x 96 bytes. The band
// depends on the block's byte phase mod 3 (blocks at 0, 32, 64 have phases 0, 2, 1): the A
// operand of a (row, phase, step) is built from the packed weight row with one v_perm per dword
// (byte selectors from a table fixed at compile time).
// Input bytes are biased (XOR 0x80) into int8 on the way to the MFMA; the accumulators start at
// 128 * sum(w), which undoes the bias, as in the dot4 engine. Tiles are 128 pixels x 32 rows, one
// wave per 32 pixels; the raw window of the next tile streams into the other of two LDS buffers
// while a tile computes. At a view edge the out-of-view pixels of the window are overwritten with
// the edge pixel (the reference clamps rows and columns to the image) before the products.
static constexpr int FM_W = 128;                  // tile width (pixels)
static constexpr int FM_H = 32;                   // tile height (rows)
static constexpr int FM_ROW_CHUNKS = 28;          // staged 16-byte chunks per row (from pixel x0 - 8)
// LDS row pitch: one unused chunk per row. The 32 lanes of a B read take the same column of 32
// rows; with a 28-chunk pitch (112 dwords) the rows fell on 2-4 bank pairs, with 29 chunks
// (58 qwords) on 16.
static constexpr int FM_PITCH_CHUNKS = 29;
static constexpr int FM_ROW = FM_PITCH_CHUNKS * 16;

template <int R>
struct FmGeom {
    static constexpr int LROWS = FM_H + 2 * R;
    static constexpr int CHUNKS = LROWS * FM_PITCH_CHUNKS;   // chunk slots, the pad slots included
};
// bytes of a window buffer (direct-to-LDS loads write only the lanes of chunks that exist)
static constexpr int FM_BUF = FmGeom<2>::CHUNKS * 16;
// Window buffers in the ring: tile t's window is loaded while tiles t - 1 and t - 2 are computed.
static constexpr int FM_RING = 3;

// Output staging: a tile's bytes (32 rows x 384) at a 400-byte row pitch, written by the waves
// from their accumulators and read back as whole 16-byte row segments, so that a store
// instruction covers 8 rows x 128 consecutive bytes instead of 8 bytes in each of 32 rows. Rows
// 16-31 hold the two 8-byte halves of every 16-byte chunk swapped: a write instruction (one dword
// of 32 rows, rows r and r + 16 in the same four banks) then touches 64 distinct banks. Two areas:
// one is read (and stored) while the next tile's products fill the other.
static constexpr int FM_OUT_PITCH = 400;
static constexpr int FM_OUT_BYTES = FM_H * FM_OUT_PITCH;

struct FmLds {
    uint8_t buf[FM_RING][FM_BUF];
    uint8_t out[2][FM_OUT_BYTES];
};

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

// Byte selectors of the A operands. Lane l holds output byte m = l % 32 of a block and the K bytes
// k = 16 (l / 32) + 0..15 of a step (the B operand uses the same K numbering for its 16 bytes).
// Byte k of step s of a block of phase p is tap dx of output m when
//   k + 32 s - 16 = m + 3 dx + 2 - 2 c,  c = (p + m) mod 3
// (input byte relative to the block's first output byte, the window starting 16 bytes before it,
// so that a window is 16-byte aligned in LDS whenever the tile's first byte is).
// Selector byte = dx + R (a byte of the packed weight row: bytes 0-3 of lo, byte 4 = byte 0 of
// hi), or 0x0c (zero).
struct FmSel {
    uint32_t v[2][64][3][2][4];   // [R - 1][lane][block][step][dword]
};
constexpr FmSel fm_make_sel() {
    FmSel t{};
    for (int r = 1; r <= 2; ++r)
        for (int l = 0; l < 64; ++l)
            for (int b = 0; b < 3; ++b)
                for (int st = 0; st < 2; ++st)
                    for (int d = 0; d < 4; ++d) {
                        uint32_t sel = 0;
                        const int m = l % 32, h = l / 32, p = (32 * b) % 3, c = (p + m) % 3;
                        for (int bb = 0; bb < 4; ++bb) {
                            const int k = 16 * h + 4 * d + bb;
                            const int u = k + 32 * st - 16 - m - 2 + 2 * c;   // 3 dx
                            uint32_t byte = 0x0c;
                            if (u >= -3 * r && u <= 3 * r && (u + 30) % 3 == 0) byte = (uint32_t)(u / 3 + r);
                            sel |= byte << (8 * bb);
                        }
                        t.v[r - 1][l][b][st][d] = sel;
                    }
    return t;
}
__device__ const FmSel g_fm_sel = fm_make_sel();

// This lane's A operands, (2R + 1) rows x 3 blocks x 2 K steps, built once per work item with one
// v_perm per dword from the packed weight row of each filter row.
template <int R>
struct FmA {
    i32x4 op[2 * R + 1][3][2];
};

template <int R>
__device__ __forceinline__ void fm_build_a(const FilterWeights& w, FmA<R>& A) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            uint32_t sel[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) sel[d] = g_fm_sel.v[R - 1][lane][b][st][d];
#pragma unroll
            for (int dy = 0; dy < 2 * R + 1; ++dy) {
                const uint32_t lo = w.wv[(dy * 4 + R) * 2], hi = w.wv[(dy * 4 + R) * 2 + 1];
#pragma unroll
                for (int d = 0; d < 4; ++d) A.op[dy][b][st][d] = (int32_t)__builtin_amdgcn_perm(hi, lo, sel[d]);
            }
        }
}

__device__ __forceinline__ void fm_origin(const FilterArgs& a, uint32_t tile, int& x0, int& y0) {
    const uint32_t tiles_y = (uint32_t)((a.rows + FM_H - 1) / FM_H);
    x0 = a.col0 + (int)(tile / tiles_y) * FM_W;
    y0 = a.row0 + (int)(tile % tiles_y) * FM_H;
}

// Offset from the view's origin of pixel x0 - 8 of the window's staged row lr (image row
// y0 - R + lr clamped to the view).
template <int R>
__device__ __forceinline__ int fm_row_start(const FilterArgs& a, const FilterView& v, int x0, int y0, int lr) {
    const int m = clampi(y0 - R + lr, v.first_row, v.last_row);
    return (int)__umul24((uint32_t)(a.view_last_row - m), (uint32_t)a.in_stride) + (x0 - 8 - v.first_col) * 3;
}

// Wave index as a scalar (the offsets built from it stay out of vector registers).
__device__ __forceinline__ int fm_wave() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Raw window of a tile: staged row lr holds the FM_ROW_CHUNKS 16-byte chunks from the 16-byte
// boundary at or below pixel x0 - 8 of image row y0 - R + lr (rows clamped to the view), at row
// pitch FM_ROW. The chunks come through a buffer resource spanning the view (chunks past its end
// read as zero, a chunk wholly before it is redirected to offset 0; see filter_dma) straight into
// LDS: wave w's load i fills chunk slots 256 i + 64 w + lane (pad slots masked off), so no
// register holds them and the ring can hold two tiles in flight.
static constexpr int FM_DMA_MAX = (FmGeom<2>::CHUNKS + FT_THREADS - 1) / FT_THREADS;   // loads per window

// This lane's part of the window loads, fixed for a work item: whether load i has a chunk for it
// and, for rows that need no clamping at a 16-byte multiple stride, the chunk's offset from the
// window's first row start (16 ch - lr stride for staged row lr, chunk ch).
struct FmDmaLane {
    int k[FM_DMA_MAX];
    bool on[FM_DMA_MAX];
};

template <int R>
__device__ __forceinline__ void fm_dma_lane(const FilterArgs& a, FmDmaLane& d) {
    constexpr int CHUNKS = FmGeom<R>::CHUNKS;
    const int wave = fm_wave(), lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < FM_DMA_MAX; ++i) {
        const int c = i * FT_THREADS + wave * 64 + lane;
        const int lr = c / FM_PITCH_CHUNKS, ch = c - lr * FM_PITCH_CHUNKS;
        d.on[i] = c < CHUNKS && ch < FM_ROW_CHUNKS;
        d.k[i] = 16 * ch - (int)__umul24((uint32_t)lr, (uint32_t)a.in_stride);
    }
}

template <int R>
__device__ __forceinline__ void fm_dma(const FilterArgs& a, const uint8_t* __restrict__ in, const FmDmaLane& d,
                                       int x0, int y0, uint8_t* __restrict__ buf) {
    constexpr int CHUNKS = FmGeom<R>::CHUNKS;
    const FilterView v = filter_view(a);
    const int mis = (int)((uintptr_t)in & 15u);
    const int extent = (int)__umul24((uint32_t)(a.view_last_row - v.first_row), (uint32_t)a.in_stride) + a.pad * 3;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(in - mis), 0, (mis + extent + 15) & ~15, 0x00020000);
    const int wave = fm_wave();
    // rows y0 - R .. y0 - R + LROWS - 1 all inside the view: one offset per window plus the lane's
    const bool inner = (a.in_stride & 15u) == 0 && y0 - R >= v.first_row &&
                       y0 - R + FmGeom<R>::LROWS - 1 <= v.last_row;
    const int base = (fm_row_start<R>(a, v, x0, y0, 0) + mis) & ~15;
#pragma unroll
    for (int i = 0; i < FM_DMA_MAX; ++i) {
        const int c0 = i * FT_THREADS + wave * 64;
        if (c0 >= CHUNKS) break;
        if (!d.on[i]) continue;
        int off;
        if (inner) {
            off = base + d.k[i];
        } else {
            int c = c0 + (int)(threadIdx.x & 63);
            asm volatile("" : "+v"(c));               // recomputed here, not kept for every tile
            const int lr = c / FM_PITCH_CHUNKS, ch = c - lr * FM_PITCH_CHUNKS;
            off = ((fm_row_start<R>(a, v, x0, y0, lr) + mis) & ~15) + ch * 16;
        }
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(buf + c0 * 16), 16,
            (uint32_t)(off < 0 ? 0 : off), 0, 0, 0);
    }
}

// Direct-to-LDS loads every wave issues per window (wave 3 has the fewest): the wait that leaves
// this many outstanding has seen the previous window land, since loads complete in order.
template <int R>
__host__ __device__ constexpr int fm_dma_min() {
    return (FmGeom<R>::CHUNKS - 3 * 64 + FT_THREADS - 1) / FT_THREADS;
}

// Byte of pixel x0 - 8 in staged row lr.
template <int R>
__device__ __forceinline__ int fm_row_byte(const FilterArgs& a, const FilterView& v, int x0, int y0, int lr, int mis) {
    return lr * FM_ROW + ((fm_row_start<R>(a, v, x0, y0, lr) + mis) & 15);
}

// Out-of-view pixels of the window that valid outputs read get the edge pixel's bytes.
template <int R>
__device__ __forceinline__ bool fm_needs_patch(const FilterArgs& a, const FilterView& v, int x0) {
    const int cend = a.col0 + a.cols;
    const int xl = (x0 + FM_W < cend ? x0 + FM_W : cend) - 1;
    return x0 - R < v.first_col || xl + R > v.last_col;
}

template <int R>
__device__ __forceinline__ void fm_patch(const FilterArgs& a, int x0, int y0, uint8_t* buf) {
    const FilterView v = filter_view(a);
    const int mis = (int)((uintptr_t)a.in & 15u);
    constexpr int LROWS = FmGeom<R>::LROWS;
    // (row, side, k): pixel first_col - 1 - k or last_col + 1 + k, k < R
    for (int t = threadIdx.x; t < LROWS * 2 * R; t += FT_THREADS) {
        const int lr = t / (2 * R), side = (t / R) & 1, k = t % R;
        const int px = side ? v.last_col + 1 + k : v.first_col - 1 - k;
        const int src = side ? v.last_col : v.first_col;
        if (px < x0 - R || px > x0 + FM_W - 1 + R) continue;
        const int rb = fm_row_byte<R>(a, v, x0, y0, lr, mis);
        uint8_t* d = buf + rb + (px - (x0 - 8)) * 3;
        const uint8_t* sp = buf + rb + (src - (x0 - 8)) * 3;
        const uint8_t b0 = sp[0], b1 = sp[1], b2 = sp[2];
        d[0] = b0; d[1] = b1; d[2] = b2;
    }
}

// 16 bytes of LDS at byte address q, as an int8 MFMA operand biased by XOR 0x80.
// AL: the alignment every q of the tile has (16, 8, or 1 for none); the 16-byte reads of a B
// operand (32 rows of one column at an odd number of chunks apart) hit distinct banks.
template <int AL>
__device__ __forceinline__ i32x4 fm_load_b(const uint8_t* base, int q) {
    uint32_t d[4];
    if constexpr (AL == 16) {
        const u32x4 u = *(const u32x4*)(base + q);
        d[0] = u.x; d[1] = u.y; d[2] = u.z; d[3] = u.w;
    } else if constexpr (AL == 8) {
        const uint2 u0 = *(const uint2*)(base + q), u1 = *(const uint2*)(base + q + 8);
        d[0] = u0.x; d[1] = u0.y; d[2] = u1.x; d[3] = u1.y;
    } else {
        const uint32_t* w = (const uint32_t*)(base + (q & ~3));
        const uint32_t sh = (uint32_t)q & 3u;
        const uint32_t e0 = w[0], e1 = w[1], e2 = w[2], e3 = w[3], e4 = w[4];
        d[0] = __builtin_amdgcn_alignbyte(e1, e0, sh);
        d[1] = __builtin_amdgcn_alignbyte(e2, e1, sh);
        d[2] = __builtin_amdgcn_alignbyte(e3, e2, sh);
        d[3] = __builtin_amdgcn_alignbyte(e4, e3, sh);
    }
    i32x4 r;
    r.x = (int32_t)(d[0] ^ 0x80808080u);
    r.y = (int32_t)(d[1] ^ 0x80808080u);
    r.z = (int32_t)(d[2] ^ 0x80808080u);
    r.w = (int32_t)(d[3] ^ 0x80808080u);
    return r;
}

// Products of one tile for this wave's 96 bytes (three 32-byte blocks x 32 rows).
struct FmAcc {
    i32x16 v[3];
};

template <int R, int AL>
__device__ __forceinline__ void fm_compute(const FilterArgs& a, const FilterWeights& w, const FmA<R>& A, int x0, int y0,
                                           const uint8_t* buf, FmAcc& acc) {
    const FilterView v = filter_view(a);
    const int mis = (int)((uintptr_t)a.in & 15u);
    const int lane = threadIdx.x & 63, wave = fm_wave();
    const int n = lane & 31, h = lane >> 5;
    const int bias = 128 * w.wsum;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc.v[b][r] = bias;
#pragma unroll
    for (int dy = 0; dy < 2 * R + 1; ++dy) {
        // tile byte 96 wave + 32 q - 16 + 16 h of staged row n + dy (pixel x0 - 8 is 24 bytes before tile byte 0)
        const int q0 = fm_row_byte<R>(a, v, x0, y0, n + dy, mis) + 8 + 96 * wave + 16 * h;
        const i32x4 b0 = fm_load_b<AL>(buf, q0);
        const i32x4 b1 = fm_load_b<AL>(buf, q0 + 32);
        const i32x4 b2 = fm_load_b<AL>(buf, q0 + 64);
        const i32x4 b3 = fm_load_b<AL>(buf, q0 + 96);
        acc.v[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][0][0], b0, acc.v[0], 0, 0, 0);
        acc.v[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][1][0], b1, acc.v[1], 0, 0, 0);
        acc.v[2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][2][0], b2, acc.v[2], 0, 0, 0);
        acc.v[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][0][1], b1, acc.v[0], 0, 0, 0);
        acc.v[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][1][1], b2, acc.v[1], 0, 0, 0);
        acc.v[2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A.op[dy][2][1], b3, acc.v[2], 0, 0, 0);
    }
}

typedef short fm_i16x2 __attribute__((ext_vector_type(2)));

// Four products as output bytes (byte 0 = e0). Mode 0 keeps the low bytes. Mode 1 takes
// clamp(s >> shift, 0, 255); for shift <= 7 (NARROW) the sums are first saturated to 16 bits in
// pairs, which gives the same bytes: a sum above 32767 shifts to at least 32767 >> 7 = 255, one
// below -32768 to a negative value, so both clamp as before.
template <int MODE, bool NARROW>
__device__ __forceinline__ uint32_t fm_word(int e0, int e1, int e2, int e3, int shift) {
    if constexpr (MODE == 0) {
        return pack4((uint32_t)e0, (uint32_t)e1, (uint32_t)e2, (uint32_t)e3);
    } else if constexpr (NARROW) {
        const fm_i16x2 sh = {(short)shift, (short)shift};
        const fm_i16x2 p01 = __builtin_amdgcn_cvt_pk_i16(e0, e1) >> sh;
        const fm_i16x2 p23 = __builtin_amdgcn_cvt_pk_i16(e2, e3) >> sh;
        uint32_t q01, q23;
        asm("v_sat_pk_u8_i16 %0, %1" : "=v"(q01) : "v"(__builtin_bit_cast(uint32_t, p01)));
        asm("v_sat_pk_u8_i16 %0, %1" : "=v"(q23) : "v"(__builtin_bit_cast(uint32_t, p23)));
        return __builtin_amdgcn_perm(q23, q01, 0x05040100u);
    } else {
        return pack4((uint32_t)clampi(e0 >> shift, 0, 255), (uint32_t)clampi(e1 >> shift, 0, 255),
                     (uint32_t)clampi(e2 >> shift, 0, 255), (uint32_t)clampi(e3 >> shift, 0, 255));
    }
}

// A tile's products as bytes into an output staging area. D[m][n]: column n = lane % 32 (tile
// row), row m = 8 (r / 4) + 4 (lane / 32) + r % 4 (byte of the block): each lane holds 4
// consecutive bytes at 12 places of its row, 8 (r / 4) bytes apart. The writes go one dword at a
// time (the empty asm statements keep the compiler from pairing them, which would bring the bank
// conflicts back) at immediate offsets from two bases, one per half of a 16-byte chunk.
template <int MODE, bool NARROW>
__device__ __forceinline__ void fm_stage_words(const FmAcc& acc, int shift, uint8_t* out) {
    const int lane = threadIdx.x & 63, wave = fm_wave();
    const int n = lane & 31, h = lane >> 5;
    const int sw = (n >> 4) * 8;                  // half swap of rows 16-31
    uint8_t* row = out + n * FM_OUT_PITCH + 96 * wave + 4 * h;
    uint8_t* half[2] = {row + sw, row + (8 ^ sw)};
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t word = fm_word<MODE, NARROW>(acc.v[b][4 * i], acc.v[b][4 * i + 1], acc.v[b][4 * i + 2],
                                                        acc.v[b][4 * i + 3], shift);
            asm volatile("" ::: "memory");
            *(uint32_t*)(half[i & 1] + 32 * b + 16 * (i >> 1)) = word;
        }
}

template <int MODE>
__device__ __forceinline__ void fm_stage_out(const FilterArgs& a, const FmAcc& acc, uint8_t* out) {
    if (MODE == 1 && a.shift <= 7) fm_stage_words<MODE, true>(acc, a.shift, out);
    else fm_stage_words<MODE, false>(acc, a.shift, out);
}

// Stores of a staged tile: thread t moves row t / 8, bytes 16 (t % 8) + 128 k, k < 3, so a
// wave's store instruction writes 8 rows x 128 consecutive bytes.
__device__ __forceinline__ void fm_store_out(const FilterArgs& a, int x0, int y0, const uint8_t* out) {
    const int r = (int)threadIdx.x >> 3, s16 = 16 * ((int)threadIdx.x & 7);
    const int y = y0 + r;
    if (y >= a.row0 + a.rows) return;
    const int nbytes = a.cols * 3;
    const int j0 = (x0 - a.col0) * 3 + s16;
    const uint8_t* src = out + r * FM_OUT_PITCH + s16;
    UH_GLOBAL uint8_t* dst = as_global(a.out) + mul24_wide((uint32_t)(y - a.row0), (uint32_t)a.out_stride) + j0;
    const bool fast = j0 + 256 + 16 <= nbytes && (((uintptr_t)dst) & 15u) == 0;
    const bool swapped = fm_wave() >= 2;          // rows 16-31 (a scalar test)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u32x4 v = *(const u32x4*)(src + 128 * k);
        if (fast) {
            if (swapped) __builtin_nontemporal_store(u32x4{v.z, v.w, v.x, v.y}, (UH_GLOBAL u32x4*)(dst + 128 * k));
            else __builtin_nontemporal_store(v, (UH_GLOBAL u32x4*)(dst + 128 * k));
        } else {
            if (swapped) v = u32x4{v.z, v.w, v.x, v.y};
            const int j = j0 + 128 * k;
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 16; ++e)
                if (j + e < nbytes) dst[128 * k + e] = (uint8_t)(w4[e >> 2] >> (8 * (e & 3)));
        }
    }
}

// Workgroup barrier that waits only for this wave's LDS accesses and for all but its last N
// vector memory operations (__syncthreads would wait for every load in flight, the window two
// tiles ahead included). The memory clobber keeps the compiler's LDS accesses on their side.
template <int N>
__device__ __forceinline__ void fm_barrier() {
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    if constexpr (N == 0) asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" :: "n"(N) : "memory");
}
__device__ __forceinline__ void fm_barrier_lds() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One step of the tile pipeline: store tile j - 3 (staged in outr), load tile j's window into
// nxt, compute tile j - 2 from cur and stage it in outw, then meet at a barrier once tile j - 1's
// window has landed. The four areas are distinct; their restrict qualifiers let the compiler see
// that the window loads in flight (into nxt) do not touch what this step reads from LDS.
template <int R, int MODE>
__device__ __forceinline__ void fm_step(const FilterArgs& a, const FilterWeights& w, const FmA<R>& A,
                                        const FmDmaLane& dl, const uint8_t* __restrict__ in,
                                        uint8_t* __restrict__ cur, uint8_t* __restrict__ nxt,
                                        uint8_t* __restrict__ outw, const uint8_t* __restrict__ outr,
                                        bool st, int sx, int sy, bool ld, int lx, int ly,
                                        bool cp, int cx, int cy) {
    if (st) fm_store_out(a, sx, sy, outr);
    if (ld) fm_dma<R>(a, in, dl, lx, ly, nxt);
    if (cp) {
        const FilterView v = filter_view(a);
        if (fm_needs_patch<R>(a, v, cx)) {
            fm_patch<R>(a, cx, cy, cur);
            fm_barrier_lds();
        }
        const int mis = (int)((uintptr_t)a.in & 15u);
        const int q = fm_row_byte<R>(a, v, cx, cy, 0, mis) + 8;   // first B byte of row 0, wave 0
        FmAcc acc;
        // products at raised wave priority: the other workgroup's wave on this SIMD issues its
        // staging, stores and loads in the MFMAs' shadow instead of ahead of them
        __builtin_amdgcn_s_setprio(1);
        if ((a.in_stride & 15u) == 0 && (q & 15) == 0) fm_compute<R, 16>(a, w, A, cx, cy, cur, acc);
        else if ((a.in_stride & 7u) == 0 && (q & 7) == 0) fm_compute<R, 8>(a, w, A, cx, cy, cur, acc);
        else fm_compute<R, 1>(a, w, A, cx, cy, cur, acc);
        __builtin_amdgcn_s_setprio(0);
        fm_stage_out<MODE>(a, acc, outw);
    }
    if (ld) fm_barrier<fm_dma_min<R>()>();
    else fm_barrier<0>();
}

// Tiles l0 .. l1-1 of one item (all tiles in one item share its arguments and weights), as steps
// j = 0 .. n + 2 of fm_step: the window of tile j goes to buf[j mod 3], so two windows are in
// flight while a tile is computed, and tile j's bytes are staged in out[j mod 2].
template <int R, int MODE>
__device__ __forceinline__ void fm_run(FmLds& sm, const FilterArgs& a, const FilterWeights& w, uint32_t l0, uint32_t l1) {
    if (l0 >= l1) return;
    FmA<R> A;
    fm_build_a<R>(w, A);
    FmDmaLane dl;
    fm_dma_lane<R>(a, dl);
    __syncthreads();              // a previous item's last step may still be reading the buffers
    const int n = (int)(l1 - l0);
    int x1 = 0, y1 = 0, x2 = 0, y2 = 0, x3 = 0, y3 = 0;   // origins of tiles j - 1, j - 2, j - 3
    int r3 = 0;                   // j mod 3
    int xn, yn;                   // origin of tile l0 + j, advanced down the column strips
    fm_origin(a, l0, xn, yn);
    const int yend = a.row0 + a.rows;
    for (int j = 0; j <= n + 2; ++j) {
        int x0 = 0, y0 = 0;
        const bool ld = j < n;
        if (ld) {
            x0 = xn; y0 = yn;
            yn += FM_H;
            if (yn >= yend) { yn = a.row0; xn += FM_W; }
        }
        const int rc = r3 == 2 ? 0 : r3 + 1;              // (j - 2) mod 3 = (j + 1) mod 3
        fm_step<R, MODE>(a, w, A, dl, a.in, sm.buf[rc], sm.buf[r3], sm.out[j & 1], sm.out[(j + 1) & 1],
                         j >= 3, x3, y3, ld, x0, y0, j >= 2 && j - 2 < n, x2, y2);
        x3 = x2; y3 = y2;
        x2 = x1; y2 = y1;
        x1 = x0; y1 = y0;
        r3 = rc;
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

// Kernel variants. Variant 0 dispatches on each item's engine, radius and mode (uniform per
// tile; the direct path stages nothing); variants 1..4 (dot4 engine) and 5..8 (MFMA engine) are
// specialised for one radius and mode (V = 1 + 2 (R - 1) + MODE, plus 4 for the MFMA engine) and
// are launched only on items of that kind, which keeps the register allocation of the common case
// free of the other paths. An item's engine (FilterWeights::pad[0], 1 = MFMA) also fixes its tile
// geometry (256 x 16 or 128 x 32 pixels), which every variant follows.
__host__ __device__ constexpr int filter_variant(int radius, int mode, int mfma) {
    return (radius == 1 || radius == 2) ? 1 + 2 * (radius - 1) + mode + (mfma ? 4 : 0) : 0;
}

// Shared memory of a variant: only variant 0 carries both engines' buffers, so the specialised
// variants keep the LDS footprint (and workgroups per CU) of their own engine.
union FilterSharedAll {
    FilterLds d;
    FmLds m;
};
struct FilterSharedDot4 {
    FilterLds d;
};
struct FilterSharedMfma {
    FmLds m;
};
template <int V>
using FilterShared = typename std::conditional<V == 0, FilterSharedAll,
                     typename std::conditional<(V >= 5), FilterSharedMfma, FilterSharedDot4>::type>::type;

template <int V>
__device__ __forceinline__ bool filter_staged(const FilterArgs& a) {
    if constexpr (V == 0) return a.radius == 1 || a.radius == 2;
    else return true;
}

template <int V>
__device__ __forceinline__ void filter_dma_v(const FilterArgs& a, int x0, int y0, uint32_t* raw) {
    if constexpr (V == 0) {
        if (a.radius == 2) filter_dma<2>(a, x0, y0, raw);
        else if (a.radius == 1) filter_dma<1>(a, x0, y0, raw);
    } else {
        filter_dma<(V + 1) / 2>(a, x0, y0, raw);
    }
}

template <int V>
__device__ __forceinline__ void filter_unpack_v(const FilterArgs& a, int x0, int y0, const uint32_t* raw, uint32_t* lds) {
    if constexpr (V == 0) {
        if (a.radius == 2) filter_unpack<2>(a, x0, y0, raw, lds);
        else if (a.radius == 1) filter_unpack<1>(a, x0, y0, raw, lds);
    } else {
        filter_unpack<(V + 1) / 2>(a, x0, y0, raw, lds);
    }
}

// Dot products of a staged tile; a tile of another radius (variant 0 only) is computed and
// stored here by the direct path. weights: the item's full weight table (direct path only).
template <int V>
__device__ __forceinline__ void filter_compute_v(const FilterArgs& a, const FilterWeights& w, const int8_t* weights,
                                                 uint32_t tile, const uint32_t* lds, FilterAcc& acc) {
    if constexpr (V == 0) {
        if (a.radius == 2) filter_accumulate<2>(w, lds, acc);
        else if (a.radius == 1) filter_accumulate<1>(w, lds, acc);
        else filter_tile_direct(a, weights, tile);
    } else {
        filter_accumulate<(V + 1) / 2>(w, lds, acc);
    }
}

template <int V>
__device__ __forceinline__ void filter_epilogue_v(const FilterArgs& a, int x0, int y0, const FilterAcc& acc) {
    if constexpr (V == 0) {
        if (!filter_staged<0>(a)) return;
        if (a.mode == 0) filter_epilogue<0>(a, x0, y0, acc);
        else filter_epilogue<1>(a, x0, y0, acc);
    } else {
        filter_epilogue<(V + 1) % 2>(a, x0, y0, acc);
    }
}


// The tile pipeline shared by both kernels: tiles l0 .. l1-1 of one item, whose arguments and
// weights stay in scalar registers. Per tile n: the dot products of n from the LDS planes, then
// (once every wave has left the planes) the de-interleave of tile n+1, whose raw window was
// requested while tile n was de-interleaved, then the request for tile n+2's window, then n's
// stores. No register holds a tile's input in flight, and the wait for a window covers stores
// issued a whole tile earlier.
template <int V>
__device__ __forceinline__ void filter_run(FilterLds& sm, const FilterArgs& a, const FilterWeights& w,
                                           const int8_t* table, uint32_t l0, uint32_t l1) {
    if (l0 >= l1) return;
    const bool staged = filter_staged<V>(a);
    // tile origins advance down the column strip: one division per segment, none per tile
    const uint32_t tiles_y = (uint32_t)((a.rows + FT_H - 1) / FT_H);
    uint32_t tx = l0 / tiles_y, ty = l0 - tx * tiles_y;
    auto origin = [&](int& x0, int& y0) {   // origin of tile (tx, ty), then step to the next tile
        x0 = a.col0 + (int)tx * FT_W;
        y0 = a.row0 + (int)ty * FT_H;
        if (++ty == tiles_y) { ty = 0; ++tx; }
    };
    int x0, y0, x1 = 0, y1 = 0, x2 = 0, y2 = 0;
    origin(x0, y0);
    if (staged) filter_dma_v<V>(a, x0, y0, sm.raw);
    wait_vm_all();
    __syncthreads();          // the window is in LDS; every wave has left the planes
    if (staged) filter_unpack_v<V>(a, x0, y0, sm.raw, sm.planes);
    __syncthreads();
    if (l0 + 1 < l1) {
        origin(x1, y1);
        if (staged) filter_dma_v<V>(a, x1, y1, sm.raw);
    }
    for (uint32_t t = l0;; ++t) {
        FilterAcc acc;
        filter_compute_v<V>(a, w, table, t, sm.planes, acc);
        const bool more = t + 1 < l1;
        if (more) {
            __syncthreads();      // every wave is done with the planes
            wait_vm_all();
            __syncthreads();      // tile t+1's window is in LDS
            if (staged) filter_unpack_v<V>(a, x1, y1, sm.raw, sm.planes);
            __syncthreads();      // planes hold tile t+1; the raw window is free
            if (t + 2 < l1) {
                origin(x2, y2);
                if (staged) filter_dma_v<V>(a, x2, y2, sm.raw);
            }
        }
        filter_epilogue_v<V>(a, x0, y0, acc);
        if (!more) break;
        x0 = x1; y0 = y1;
        x1 = x2; y1 = y2;
    }
}

// Plain launch: workgroup b computes tiles 2b and 2b + 1 of one item.
// Tiles l0 .. l1-1 of one item with the engine its variant (or, for variant 0, the item) names.
template <int V>
__device__ __forceinline__ void filter_item_run(FilterShared<V>& sm, const FilterArgs& a, const FilterWeights& w,
                                                const int8_t* table, uint32_t l0, uint32_t l1) {
    if constexpr (V >= 5) {
        fm_run<(V - 3) / 2, (V - 5) % 2>(sm.m, a, w, l0, l1);
    } else if constexpr (V == 0) {
        if (w.pad[0] == 1 && (a.radius == 1 || a.radius == 2)) {
            if (a.radius == 1) {
                if (a.mode == 0) fm_run<1, 0>(sm.m, a, w, l0, l1);
                else fm_run<1, 1>(sm.m, a, w, l0, l1);
            } else {
                if (a.mode == 0) fm_run<2, 0>(sm.m, a, w, l0, l1);
                else fm_run<2, 1>(sm.m, a, w, l0, l1);
            }
        } else {
            filter_run<0>(sm.d, a, w, table, l0, l1);
        }
    } else {
        filter_run<V>(sm.d, a, w, table, l0, l1);
    }
}

// Waves per SIMD the variant is compiled for: the MFMA engine (also reachable from variant 0)
// keeps its A operands in registers.
template <int V>
struct FilterOcc {
    static constexpr int value = (V == 0 || V >= 5) ? 2 : FT_WAVES_PER_EU;
};

template <int V>
__global__ void __launch_bounds__(FT_THREADS) __attribute__((amdgpu_waves_per_eu(FilterOcc<V>::value)))
k_imagefilter(FilterItem it, uint32_t tiles) {
    __shared__ FilterShared<V> sm;
    const uint32_t l0 = blockIdx.x * 2u;
    filter_item_run<V>(sm, head_of(it.a), it.w, it.a.weights, l0, l0 + 2u < tiles ? l0 + 2u : tiles);
}

// Persistent launch over a device work queue with a static schedule: worker w runs a contiguous
// range of the queue's tile sequence (no per-tile atomics: a single counter hit by every
// workgroup for every tile serialises across the XCDs), in segments that each lie in one item.
// Workgroups are numbered so that the workers of one XCD (hardware workgroup b runs on XCD
// b mod 8) are consecutive: neighbouring strips share an L2. Every workgroup leaves after its
// range, so the grid always drains.
static constexpr uint32_t FT_PREFIX_LDS = 1024;   // items whose tile prefix is searched in LDS

template <int V>
__global__ void __launch_bounds__(FT_THREADS) __attribute__((amdgpu_waves_per_eu(FilterOcc<V>::value)))
k_imagefilter_persistent(const uint8_t* queue) {
    __shared__ FilterShared<V> sm;
    __shared__ uint32_t s_prefix[FT_PREFIX_LDS + 1];
    const QueueHeader hd = uniform_load((const QueueHeader*)queue);   // n_items, total_tiles, offsets
    const FilterItem* items = (const FilterItem*)(queue + hd.items_offset);
    const uint32_t* prefix = (const uint32_t*)(queue + hd.prefix_offset);
    const bool lds_prefix = stage_prefix(prefix, hd.n_items, s_prefix, FT_PREFIX_LDS + 1);
    __syncthreads();
    const uint32_t g = gridDim.x, b = blockIdx.x, total = hd.total_tiles;
    const uint32_t worker = (g % 8u == 0u) ? (b % 8u) * (g / 8u) + b / 8u : b;
    const uint32_t base = total / g, rem = total % g;
    uint32_t t = worker * base + (worker < rem ? worker : rem);
    const uint32_t end = t + base + (worker < rem ? 1u : 0u);
    if (t >= end) return;
    uint32_t item, local;
    if (lds_prefix) queue_locate_lds(s_prefix, hd.n_items, t, item, local);
    else queue_locate(prefix, hd.n_items, t, item, local);
    while (t < end) {
        const uint32_t item_end = uniform_load(prefix + item + 1);   // first tile of the next item
        const uint32_t n = (end < item_end ? end : item_end) - t;
        if (n) {
            const FilterArgs a = uniform_load((const FilterArgs*)&items[item].a);
            const FilterWeights w = uniform_load(&items[item].w);
            filter_item_run<V>(sm, a, w, items[item].a.weights, local, local + n);
        }
        t += n;
        ++item;
        local = 0;
    }
}

uint32_t filter_tiles(const FilterItem& it) {
    if (it.w.pad[0] == 1) return (uint32_t)(((it.a.cols + FM_W - 1) / FM_W) * ((it.a.rows + FM_H - 1) / FM_H));
    return (uint32_t)(((it.a.cols + FT_W - 1) / FT_W) * ((it.a.rows + FT_H - 1) / FT_H));
}

int filter_variant_for(uint32_t kinds) {
    // one specialised variant present and nothing else
    return (kinds && (kinds & (kinds - 1)) == 0 && !(kinds & 1u)) ? __builtin_ctz(kinds) : 0;
}

static constexpr int FILTER_VARIANTS = 9;
typedef void (*FilterPersistentFn)(const uint8_t*);
static const FilterPersistentFn kFilterPersistent[FILTER_VARIANTS] = {
    k_imagefilter_persistent<0>, k_imagefilter_persistent<1>, k_imagefilter_persistent<2>,
    k_imagefilter_persistent<3>, k_imagefilter_persistent<4>, k_imagefilter_persistent<5>,
    k_imagefilter_persistent<6>, k_imagefilter_persistent<7>, k_imagefilter_persistent<8>};
typedef void (*FilterPlainFn)(FilterItem, uint32_t);
static const FilterPlainFn kFilterPlain[FILTER_VARIANTS] = {
    k_imagefilter<0>, k_imagefilter<1>, k_imagefilter<2>, k_imagefilter<3>, k_imagefilter<4>,
    k_imagefilter<5>, k_imagefilter<6>, k_imagefilter<7>, k_imagefilter<8>};

// Workgroups of a persistent filter kernel resident per CU (occupancy queries, cached).
static int g_filter_wg[FILTER_VARIANTS] = {};

int filter_wg_per_cu() {
    if (!g_filter_wg[0]) {
        for (int v = 0; v < FILTER_VARIANTS; ++v) {
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
    // the staged path addresses the view through a buffer resource with 32-bit offsets
    const int fr = a->row0 - r < 0 ? 0 : a->row0 - r;
    if ((r == 1 || r == 2) && (uint64_t)(lr - fr) * a->in_stride + (uint64_t)(lc - fc + 1) * 3 + 32 >= (1ull << 31))
        return fail(UH_ERR_ARG, "uh_imagefilter: the view spans 2 GiB or more");
    // ... and computes row offsets with 24-bit multiplies
    if ((r == 1 || r == 2) && (a->in_stride >= (1u << 24) || a->out_stride >= (1u << 24) || a->height >= (1 << 24)))
        return fail(UH_ERR_ARG, "uh_imagefilter: row strides of 16 MiB or more");
    // The weighted sum must fit int32 for mode 1 (always true for |w| <= 127, r <= 15).
    return UH_OK;
}

// Engine of the staged radii: 1 = MFMA (default), 0 = dot4; UH_FILTER_ENGINE=dot4 selects the
// dot4 engine at start.
static std::atomic<int> g_filter_engine{-1};

static int filter_engine() {
    int e = g_filter_engine.load();
    if (e < 0) {
        const char* s = getenv("UH_FILTER_ENGINE");
        e = (s && strcmp(s, "dot4") == 0) ? 0 : 1;
        g_filter_engine.store(e);
    }
    return e;
}

extern "C" int uh_imagefilter_engine(int engine, int* previous) {
    if (engine < -1 || engine > 1) return fail(UH_ERR_ARG, "uh_imagefilter_engine: engine %d", engine);
    const int prev = filter_engine();
    if (previous) *previous = prev;
    if (engine >= 0) g_filter_engine.store(engine);
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
    const bool staged = it.a.radius == 1 || it.a.radius == 2;
    if (staged) pack_filter_weights(it.a.weights, it.a.radius, &it.w);
    const int mfma = staged && filter_engine() == 1;
    it.w.pad[0] = mfma;
    const uint32_t tiles = filter_tiles(it);
    const int v = filter_variant(it.a.radius, it.a.mode, mfma);
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
