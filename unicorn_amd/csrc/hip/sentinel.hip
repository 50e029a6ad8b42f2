// Sentinel compression of a reduction output before it crosses PCIe (replaces
// pmCudaInterface::CompressForSentinel, pmCudaInterface.cu:380-472, which the reference's GPU
// stub calls from pmExecutionStub.cpp:3850-3881 when a reduction subtask's output is one large
// mostly-zero address space).
//
// The compressed buffer is internal to the runtime: a uint64 count, then the count non-zero
// elements in index order (at byte 8, so 8-byte elements stay aligned), then their uint32
// indices. An element is dropped when it compares equal to zero of its type, as the reference's
// `x != (T)0` test does (so a float -0.0 comes back as +0.0); decompression zero-fills.
//
// gfx950 design: a stream compaction in three launches. Block b owns elements
// [4096 b, 4096 b + 4096), read as 16 coalesced passes of 256; (1) each block counts its
// non-zero elements, (2) one workgroup turns the counts into exclusive offsets and a total,
// (3) when the caller's bound holds, each block writes its elements at offset + running rank,
// the rank within a pass coming from wave ballots and the waves' counts in LDS. HBM-bound: the
// input is read twice when it compresses, once when it does not.
#include "uh_internal.h"

#include <map>
#include <mutex>

namespace uh {

static constexpr int SC_THREADS = 256;
static constexpr int SC_PASSES = 16;
static constexpr int SC_CHUNK = SC_THREADS * SC_PASSES;
static constexpr int SC_SCAN_THREADS = 1024;

template <typename T>
__device__ __forceinline__ bool sc_nonzero(T x) {
    return x != (T)0;
}

__device__ __forceinline__ uint32_t sc_block_sum(uint32_t v, uint32_t* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += lds[w];
    return t;
}

template <typename T>
__global__ void __launch_bounds__(SC_THREADS) k_sc_count(const T* in, uint64_t n, uint32_t* counts) {
    __shared__ uint32_t lds[SC_THREADS / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SC_CHUNK;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < SC_PASSES; ++k) {
        const uint64_t i = base + (uint64_t)(k * SC_THREADS) + threadIdx.x;
        if (i < n && sc_nonzero(in[i])) ++c;
    }
    const uint32_t t = sc_block_sum(c, lds);
    if (threadIdx.x == 0) counts[blockIdx.x] = t;
}

// Exclusive offsets of the block counts and their total, in one workgroup: thread t sums a
// contiguous run of counts, the runs' sums are scanned in LDS, and each thread writes its run.
__global__ void __launch_bounds__(SC_SCAN_THREADS) k_sc_scan(const uint32_t* counts, uint32_t blocks,
                                                             uint32_t* offsets, uint64_t* total) {
    __shared__ uint64_t part[SC_SCAN_THREADS];
    const uint32_t per = (blocks + SC_SCAN_THREADS - 1) / SC_SCAN_THREADS;
    const uint32_t b0 = threadIdx.x * per;
    const uint32_t b1 = b0 + per < blocks ? b0 + per : blocks;
    uint64_t s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += counts[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < SC_SCAN_THREADS; o <<= 1) {
        const uint64_t add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint64_t run = part[threadIdx.x] - s;     // exclusive
    for (uint32_t b = b0; b < b1; ++b) {
        offsets[b] = (uint32_t)run;
        run += counts[b];
    }
    if (threadIdx.x == SC_SCAN_THREADS - 1) *total = part[threadIdx.x];
}

template <typename T>
__global__ void __launch_bounds__(SC_THREADS) k_sc_write(const T* in, uint64_t n, const uint32_t* offsets,
                                                         const uint64_t* total, uint8_t* out) {
    __shared__ uint32_t wcnt[SC_THREADS / 64];
    const uint64_t count = *total;
    T* vals = (T*)(out + 8);
    uint32_t* idx = (uint32_t*)(out + 8 + count * sizeof(T));
    if (blockIdx.x == 0 && threadIdx.x == 0) *(uint64_t*)out = count;
    const uint64_t base = (uint64_t)blockIdx.x * SC_CHUNK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t run = offsets[blockIdx.x];
    for (int k = 0; k < SC_PASSES; ++k) {
        const uint64_t i = base + (uint64_t)(k * SC_THREADS) + threadIdx.x;
        T v = (T)0;
        const bool nz = i < n && sc_nonzero(v = in[i]);
        const uint64_t ball = __ballot(nz);
        const uint32_t rank = (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(ball);
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int w = 0; w < SC_THREADS / 64; ++w) {
            before += w < wave ? wcnt[w] : 0u;
            all += wcnt[w];
        }
        if (nz) {
            const uint32_t pos = run + before + rank;
            vals[pos] = v;
            idx[pos] = (uint32_t)i;
        }
        run += all;
        __syncthreads();                            // wcnt reuse
    }
}

// Work buffers: block counts, offsets and the total, per (device, stream), grown on demand and
// kept while the process runs.
static std::mutex g_sc_mtx;
static std::map<std::pair<int, hipStream_t>, std::pair<uint8_t*, size_t>> g_sc;

static uint8_t* sc_buffer(hipStream_t stream, size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(g_sc_mtx);
    auto& e = g_sc[std::make_pair(dev, stream)];
    if (e.second < bytes) {
        if (e.first && hipFree(e.first) != hipSuccess) return nullptr;
        e.first = nullptr;
        e.second = 0;
        if (hipMalloc((void**)&e.first, bytes) != hipSuccess) return nullptr;
        e.second = bytes;
    }
    return e.first;
}

template <typename T>
static int sc_run(const void* data, uint64_t n, void* out, uint64_t max_nonzero, uint64_t* nonzero,
                  hipStream_t s) {
    const uint32_t blocks = (uint32_t)((n + SC_CHUNK - 1) / SC_CHUNK);
    uint8_t* w = sc_buffer(s, 8 + (size_t)blocks * 8);
    if (!w) return fail(UH_ERR_HIP, "uh_sentinel_compress: work buffer");
    uint64_t* total = (uint64_t*)w;
    uint32_t* counts = (uint32_t*)(w + 8);
    uint32_t* offsets = counts + blocks;
    UH_LAUNCH(k_sc_count<T>, dim3(blocks), dim3(SC_THREADS), 0, s, (const T*)data, n, counts);
    UH_LAUNCH_CHECK("k_sc_count");
    UH_LAUNCH(k_sc_scan, dim3(1), dim3(SC_SCAN_THREADS), 0, s, (const uint32_t*)counts, blocks, offsets, total);
    UH_LAUNCH_CHECK("k_sc_scan");
    uint64_t h = 0;
    UH_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, s));
    UH_HIP(hipStreamSynchronize(s));
    g_plain_launches += 2;
    *nonzero = h;
    if (h > max_nonzero) return UH_OK;       // not compressed; out untouched
    UH_LAUNCH(k_sc_write<T>, dim3(blocks), dim3(SC_THREADS), 0, s, (const T*)data, n, (const uint32_t*)offsets,
              (const uint64_t*)total, (uint8_t*)out);
    UH_LAUNCH_CHECK("k_sc_write");
    g_plain_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_sentinel_compress(const void* data, uint64_t elems, int type, void* out, uint64_t max_nonzero,
                                    uint64_t* nonzero, void* stream) {
    if (!nonzero) return fail(UH_ERR_ARG, "uh_sentinel_compress: null count");
    *nonzero = 0;
    if (!elems) {
        if (!out) return fail(UH_ERR_ARG, "uh_sentinel_compress: null output");
        return UH_OK;
    }
    if (!data || !out) return fail(UH_ERR_ARG, "uh_sentinel_compress: null buffer");
    if (elems > 0xFFFFFFFFull) return fail(UH_ERR_ARG, "uh_sentinel_compress: %llu elements (indices are 32-bit)",
                                           (unsigned long long)elems);
    if (((uintptr_t)out & 7u) != 0) return fail(UH_ERR_ARG, "uh_sentinel_compress: output must be 8-byte aligned");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    switch (type) {
        case 0: return sc_run<int32_t>(data, elems, out, max_nonzero, nonzero, s);
        case 1: return sc_run<uint32_t>(data, elems, out, max_nonzero, nonzero, s);
        case 2: return sc_run<int64_t>(data, elems, out, max_nonzero, nonzero, s);
        case 3: return sc_run<uint64_t>(data, elems, out, max_nonzero, nonzero, s);
        case 4: return sc_run<float>(data, elems, out, max_nonzero, nonzero, s);
        case 5: return sc_run<double>(data, elems, out, max_nonzero, nonzero, s);
        default: return fail(UH_ERR_ARG, "uh_sentinel_compress: element type %d", type);
    }
}
