// Block transpose of 32-bit elements: out[c * ld_out + r] = in[r * ld_in + c] for a rows x cols
// block. Replaces matrixTranspose_cuda (reference matrixTransposeCUDA.cu:46-69), which moves
// 32x32 tiles through shared memory.
//
// gfx950 design: one workgroup of 256 threads (4 waves) moves a 64 x 64 tile. Each thread
// loads four 16-byte row segments (rows t/16 + 16k, columns 4*(t%16)..+3), so a wave reads
// 4 rows x 256 B contiguously; the tile goes through LDS with a row pitch of 65 words, which
// puts the column reads of the write phase on distinct banks; each thread then stores four
// 16-byte segments of the output rows. Tiles on the block's edge, and blocks whose rows are not
// 16-byte aligned, take the same path with element-wise guarded accesses. HBM-bound: 8 bytes move
// per element (one read, one write).
#include "uh_internal.h"

namespace uh {

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

static constexpr int TT = 64;            // tile edge
static constexpr int TP = TT + 1;        // LDS row pitch in words

__global__ void __launch_bounds__(256) k_transpose32(const uint32_t* __restrict__ in, uint64_t ld_in,
                                                      uint32_t* __restrict__ out, uint64_t ld_out, uint32_t rows,
                                                      uint32_t cols, uint32_t tiles_c, int vec_ok) {
    __shared__ uint32_t tile[TT * TP];
    const uint32_t tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
    const uint32_t r0 = tr * TT, c0 = tc * TT;
    const int t = threadIdx.x;
    const int lr = t >> 4, lc = (t & 15) * 4;
    const bool full = vec_ok && r0 + TT <= rows && c0 + TT <= cols;
    const UH_GLOBAL uint32_t* src = as_global(in);
    UH_GLOBAL uint32_t* dst = as_global(out);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t r = lr + 16 * k;
        if (full) {
            const u32x4_t v = *(const UH_GLOBAL u32x4_t*)(src + (uint64_t)(r0 + r) * ld_in + c0 + lc);
            tile[r * TP + lc + 0] = v.x;
            tile[r * TP + lc + 1] = v.y;
            tile[r * TP + lc + 2] = v.z;
            tile[r * TP + lc + 3] = v.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t gr = r0 + r, gc = c0 + lc + e;
                tile[r * TP + lc + e] = (gr < rows && gc < cols) ? src[(uint64_t)gr * ld_in + gc] : 0u;
            }
        }
    }
    __syncthreads();
    // output row = input column c0 + r, output column = input row r0 + lc + e
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t r = lr + 16 * k;
        if (full) {
            u32x4_t v;
            v.x = tile[(lc + 0) * TP + r];
            v.y = tile[(lc + 1) * TP + r];
            v.z = tile[(lc + 2) * TP + r];
            v.w = tile[(lc + 3) * TP + r];
            *(UH_GLOBAL u32x4_t*)(dst + (uint64_t)(c0 + r) * ld_out + r0 + lc) = v;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t orow = c0 + r, ocol = r0 + lc + e;
                if (orow < cols && ocol < rows) dst[(uint64_t)orow * ld_out + ocol] = tile[(lc + e) * TP + r];
            }
        }
    }
}

}  // namespace uh

using namespace uh;

extern "C" int uh_transpose32(const void* in, uint64_t ld_in, void* out, uint64_t ld_out, uint32_t rows, uint32_t cols,
                              void* stream) {
    if (!rows || !cols) return UH_OK;
    if (!in || !out) return fail(UH_ERR_ARG, "uh_transpose32: null buffer");
    if (ld_in < cols || ld_out < rows) return fail(UH_ERR_ARG, "uh_transpose32: leading dimension below the block");
    if (in == out) return fail(UH_ERR_ARG, "uh_transpose32: in and out must not alias");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    const uint64_t tiles_r = (rows + TT - 1) / TT, tiles_c = (cols + TT - 1) / TT;
    if (tiles_r * tiles_c >= (1ull << 24)) return fail(UH_ERR_ARG, "uh_transpose32: block too large");
    const int vec_ok = ((((uintptr_t)in | (uintptr_t)out) & 15u) == 0 && (ld_in & 3) == 0 && (ld_out & 3) == 0) ? 1 : 0;
    UH_LAUNCH(k_transpose32, dim3((unsigned)(tiles_r * tiles_c)), dim3(256), 0, s, (const uint32_t*)in, ld_in,
              (uint32_t*)out, ld_out, rows, cols, (uint32_t)tiles_c, vec_ok);
    UH_LAUNCH_CHECK("k_transpose32");
    g_plain_launches++;
    return UH_OK;
}
