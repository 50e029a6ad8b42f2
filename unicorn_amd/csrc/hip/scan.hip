// prefixSum kernels (replace prefixSum_cuda / elemAdd_cuda, reference prefixSumCUDA.cu:68-165):
// an inclusive scan of 32-bit integers with wrap-around arithmetic, and the addition of one
// device-resident value to every element of a range.
//
// gfx950 design. The scan is two passes over 16 KB blocks (256 threads x 16 elements): pass 1
// writes each block's total; pass 2 has every block add up the totals of the blocks before it
// (at most a few hundred values, read once per block from L2), then scans its own 16 elements
// per thread in registers, across the wave with 6 steps of __shfl_up over 64 lanes, and across
// the 4 waves through LDS. Integer addition modulo 2^32 is
// associative, so the result equals the serial left-to-right sum bit for bit. 12 bytes of HBM
// traffic per element (read twice, written once); each block's loads are 16-byte vectors.
// Block totals live in small buffers the library owns, one per stream (grown on demand, kept
// while the process runs).
#include "uh_internal.h"

#include <map>
#include <mutex>

namespace uh {

static constexpr int SC_THREADS = 256;
static constexpr int SC_PER = 16;                        // elements per thread
static constexpr int SC_BLOCK = SC_THREADS * SC_PER;     // 4096 elements per block

typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load16(const int32_t* in, size_t n, size_t base, int32_t* v, bool vec) {
    if (vec && base + SC_PER <= n) {
        const UH_GLOBAL i32x4_t* p = (const UH_GLOBAL i32x4_t*)(as_global(in) + base);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const i32x4_t x = p[q];
            v[4 * q + 0] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < SC_PER; ++e) v[e] = base + e < n ? in[base + e] : 0;
    }
}

// wave-wide inclusive scan of one value per lane (wrap-around)
__device__ __forceinline__ int32_t wave_scan(int32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x = (int32_t)((uint32_t)x + (uint32_t)y);
    }
    return x;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_totals(const int32_t* in, size_t n, int32_t* totals, int vec) {
    __shared__ int32_t wsum[SC_THREADS / 64];
    const size_t base = (size_t)blockIdx.x * SC_BLOCK + (size_t)threadIdx.x * SC_PER;
    int32_t v[SC_PER];
    load16(in, n, base, v, vec != 0);
    uint32_t s = 0;
#pragma unroll
    for (int e = 0; e < SC_PER; ++e) s += (uint32_t)v[e];
    const int32_t w = wave_scan((int32_t)s);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < SC_THREADS / 64; ++k) t += (uint32_t)wsum[k];
        totals[blockIdx.x] = (int32_t)t;
    }
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_blocks(const int32_t* in, int32_t* out, size_t n,
                                                             const int32_t* totals, int vec) {
    __shared__ int32_t wsum[SC_THREADS / 64];
    __shared__ int32_t carry_in;
    // sum of the totals of the blocks before this one
    uint32_t c = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += SC_THREADS) c += (uint32_t)totals[b];
    c = (uint32_t)wave_scan((int32_t)c);                 // lane 63 holds the wave's sum
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = (int32_t)c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < SC_THREADS / 64; ++k) t += (uint32_t)wsum[k];
        carry_in = (int32_t)t;
    }
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * SC_BLOCK + (size_t)threadIdx.x * SC_PER;
    int32_t v[SC_PER];
    load16(in, n, base, v, vec != 0);
    uint32_t s = 0;
#pragma unroll
    for (int e = 0; e < SC_PER; ++e) { s += (uint32_t)v[e]; v[e] = (int32_t)s; }   // thread-local inclusive
    const uint32_t w = (uint32_t)wave_scan((int32_t)s);                              // inclusive over lanes
    __syncthreads();                                                                 // wsum reuse
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = (int32_t)w;
    __syncthreads();
    uint32_t before = (uint32_t)carry_in + (w - s);                                  // exclusive in wave
    for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) before += (uint32_t)wsum[k];
    if (vec && base + SC_PER <= n) {
        UH_GLOBAL i32x4_t* p = (UH_GLOBAL i32x4_t*)(as_global(out) + base);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            i32x4_t x;
            x.x = (int32_t)(before + (uint32_t)v[4 * q + 0]);
            x.y = (int32_t)(before + (uint32_t)v[4 * q + 1]);
            x.z = (int32_t)(before + (uint32_t)v[4 * q + 2]);
            x.w = (int32_t)(before + (uint32_t)v[4 * q + 3]);
            p[q] = x;
        }
    } else {
#pragma unroll
        for (int e = 0; e < SC_PER; ++e)
            if (base + e < n) out[base + e] = (int32_t)(before + (uint32_t)v[e]);
    }
}

__global__ void __launch_bounds__(256) k_add_scalar_i32(int32_t* data, size_t n, const int32_t* value) {
    const uint32_t a = (uint32_t)*value;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        data[i] = (int32_t)((uint32_t)data[i] + a);
}

// Block totals: one buffer per (device, stream), so that scans on different streams never share
// one; calls on one stream are ordered, and a buffer is replaced only through hipFree, which waits
// for the device.
static std::mutex g_tot_mtx;
static std::map<std::pair<int, hipStream_t>, std::pair<int32_t*, size_t>> g_tot;

static int32_t* totals_buffer(hipStream_t stream, size_t blocks) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(g_tot_mtx);
    auto& e = g_tot[std::make_pair(dev, stream)];
    if (e.second < blocks) {
        if (e.first && hipFree(e.first) != hipSuccess) return nullptr;
        e.first = nullptr;
        e.second = 0;
        const size_t cap = blocks < 4096 ? 4096 : blocks;
        if (hipMalloc((void**)&e.first, cap * 4) != hipSuccess) return nullptr;
        e.second = cap;
    }
    return e.first;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_scan_i32(const int32_t* in, int32_t* out, uint64_t n, void* stream) {
    if (!n) return UH_OK;
    if (!in || !out) return fail(UH_ERR_ARG, "uh_scan_i32: null buffer");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    const uint64_t blocks = (n + SC_BLOCK - 1) / SC_BLOCK;
    if (blocks > 0x7fffffffull) return fail(UH_ERR_ARG, "uh_scan_i32: too many elements");
    // the block totals are read by every later block: the pass-2 cost grows with blocks^2 / 256
    if (blocks > 4096) return fail(UH_ERR_ARG, "uh_scan_i32: more than 2^24 elements per call");
    int32_t* tot = totals_buffer(s, blocks);
    if (!tot) return fail(UH_ERR_OOM, "uh_scan_i32: block totals");
    const int vec = ((((uintptr_t)in | (uintptr_t)out) & 15u) == 0) ? 1 : 0;
    UH_LAUNCH(k_scan_totals, dim3((unsigned)blocks), dim3(SC_THREADS), 0, s, in, (size_t)n, tot, vec);
    UH_LAUNCH_CHECK("k_scan_totals");
    UH_LAUNCH(k_scan_blocks, dim3((unsigned)blocks), dim3(SC_THREADS), 0, s, in, out, (size_t)n, (const int32_t*)tot, vec);
    UH_LAUNCH_CHECK("k_scan_blocks");
    g_plain_launches += 2;
    return UH_OK;
}

extern "C" int uh_add_scalar_i32(int32_t* data, uint64_t n, const int32_t* value, void* stream) {
    if (!n) return UH_OK;
    if (!data || !value) return fail(UH_ERR_ARG, "uh_add_scalar_i32: null buffer");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    uint64_t blocks = (n + 255) / 256;
    const uint64_t cap = (uint64_t)cu_count() * 8;
    if (blocks > cap) blocks = cap;
    UH_LAUNCH(k_add_scalar_i32, dim3((unsigned)blocks), dim3(256), 0, s, data, (size_t)n, value);
    UH_LAUNCH_CHECK("k_add_scalar_i32");
    g_plain_launches++;
    return UH_OK;
}
