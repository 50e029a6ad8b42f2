// Test support (built by unicorn_amd.build.build_test_support into tests/native/libreduce_probe.so):
//   rp_reduce: a task whose subtasks all write the whole output space and whose reduction is a
//     built-in one (pmGetSubtaskReductionCallbackImpl(op, type)), with a CPU callback and a
//     kernel-kind GPU callback writing the same values; the runtime folds the subtasks' private
//     outputs on the host (CPU stubs, PLAIN HIP mode) or in HBM (RESIDENT / BATCHED).
//   rp_redistribute: a task whose dataRedistribution callback cuts each subtask's block into
//     chunks and gives chunk j the order (subtask + j) % orders; the chunk lengths travel from
//     the subtask callback to the redistribution callback in a SUBTASK_TO_POST_SUBTASK scratch
//     buffer (written by the kernel on the GPU), as the reference's inverted-index app does
//     (pageRankInvertedIndex.cpp:300-395).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "pmPublicDefinitions.h"
#include "pmPublicUtilities.h"
#include "pmExtensions.h"

using namespace pm;

namespace {

// ---------------------------------------------------------------- reductions

struct RConf {
    uint32_t n;      // elements
    int op, type;
    uint32_t sparse; // nonzero: each subtask sets only about one element in 16 (the rest 0)
};

__host__ __device__ inline uint32_t mix(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

// small integers that every element type holds exactly: products of at most 8 subtasks stay
// below 2^8, sums below 8 * 3
__host__ __device__ inline long long gen(int op, int type, uint32_t sid, uint32_t i, uint32_t sparse) {
    if (sparse && (mix(sid + 101, i) & 15u) != 0) return 0;
    const uint32_t h = mix(sid + 1, i);
    const bool is_unsigned = type == REDUCE_UNSIGNED_INTS || type == REDUCE_UNSIGNED_LONGS;
    switch (op) {
        case REDUCE_PRODUCT: return (long long)(h % 2 + 1);
        case REDUCE_ADD: case REDUCE_MIN: case REDUCE_MAX:
            return is_unsigned ? (long long)(h % 7) : (long long)(h % 7) - 3;
        case REDUCE_LOGICAL_XOR: return (long long)(h % 2);   // (a != b) folds in any order on 0/1
        default: return (long long)(h % 5);    // logical ops see 0 and several true values
    }
}

template <typename T>
__global__ void rp_kernel(pmTaskInfo ti, pmDeviceInfo*, pmSubtaskInfo si, pmStatus*) {
    const RConf* c = (const RConf*)ti.taskConf;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c->n) return;
    ((T*)si.memInfo[0].writePtr)[i] = (T)gen(c->op, c->type, (uint32_t)si.subtaskId, i, c->sparse);
}

template <typename T>
void fill_cpu(const RConf* c, uint32_t sid, void* p) {
    for (uint32_t i = 0; i < c->n; ++i) ((T*)p)[i] = (T)gen(c->op, c->type, sid, i, c->sparse);
}

size_t type_size(int type) { return type == REDUCE_INTS || type == REDUCE_UNSIGNED_INTS || type == REDUCE_FLOATS ? 4 : 8; }

pmStatus r_dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const RConf* c = (const RConf*)ti.taskConf;
    pmStatus s = pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION,
                                     pmSubscriptionInfo(0, (size_t)c->n * type_size(c->type)));
    if (s == pmSuccess && dev.deviceType == GPU_HIP) {
        pmGpuLaunchConf lc((int)((c->n + 255) / 256), 1, 1, 256, 1, 1);
        s = pmSetGpuLaunchConf(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, lc);
    }
    return s;
}

pmStatus r_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const RConf* c = (const RConf*)ti.taskConf;
    void* p = si.memInfo[0].writePtr;
    switch (c->type) {
        case REDUCE_INTS: fill_cpu<int>(c, (uint32_t)si.subtaskId, p); break;
        case REDUCE_UNSIGNED_INTS: fill_cpu<unsigned int>(c, (uint32_t)si.subtaskId, p); break;
        case REDUCE_LONGS: fill_cpu<long>(c, (uint32_t)si.subtaskId, p); break;
        case REDUCE_UNSIGNED_LONGS: fill_cpu<unsigned long>(c, (uint32_t)si.subtaskId, p); break;
        case REDUCE_FLOATS: fill_cpu<float>(c, (uint32_t)si.subtaskId, p); break;
        default: fill_cpu<double>(c, (uint32_t)si.subtaskId, p); break;
    }
    return pmSuccess;
}

std::mutex g_mtx;
pmCallbackHandle g_rcb[MAX_REDUCTION_OP_TYPES][MAX_REDUCTION_DATA_TYPES][3];

pmSubtaskCallback_GPU_HIP kernel_of(int type) {
    switch (type) {
        case REDUCE_INTS: return rp_kernel<int>;
        case REDUCE_UNSIGNED_INTS: return rp_kernel<unsigned int>;
        case REDUCE_LONGS: return rp_kernel<long>;
        case REDUCE_UNSIGNED_LONGS: return rp_kernel<unsigned long>;
        case REDUCE_FLOATS: return rp_kernel<float>;
        default: return rp_kernel<double>;
    }
}

// ---------------------------------------------------------------- redistribution

struct DConf {
    uint32_t n;        // elements per subtask block
    uint32_t orders;   // chunks per block, each its own order
};

// chunk j of subtask s's block: lengths cut by a hash, summing to n
__host__ __device__ inline void chunk_lengths(const DConf* c, uint32_t sid, uint32_t* len) {
    uint32_t left = c->n;
    for (uint32_t j = 0; j < c->orders; ++j) {
        uint32_t l = j + 1 == c->orders ? left : (left ? mix(sid, j) % (left + 1) : 0);
        len[j] = l;
        left -= l;
    }
}

__global__ void dp_kernel(pmTaskInfo ti, pmDeviceInfo* dev, pmSubtaskInfo si, pmStatus*) {
    const DConf* c = (const DConf*)ti.taskConf;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c->n) ((uint32_t*)si.memInfo[0].writePtr)[i] = (uint32_t)si.subtaskId * c->n + i;
    if (i == 0) {
        uint32_t* sc = (uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev->deviceHandle, si.subtaskId, si.splitInfo,
                                                     SUBTASK_TO_POST_SUBTASK, 0, &si.gpuContext);
        chunk_lengths(c, (uint32_t)si.subtaskId, sc);
    }
}

pmStatus d_dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const DConf* c = (const DConf*)ti.taskConf;
    pmStatus s = pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION,
                                     pmSubscriptionInfo((size_t)si.subtaskId * c->n * 4, (size_t)c->n * 4));
    if (s != pmSuccess) return s;
    if (dev.deviceType == GPU_HIP) {
        // the device callback cannot allocate: the buffer is sized here and copied back after it
        if (!pmGetScratchBuffer(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, SUBTASK_TO_POST_SUBTASK,
                                (size_t)c->orders * 4, NULL))
            return pmMemoryError;
        pmGpuLaunchConf lc((int)((c->n + 255) / 256), 1, 1, 256, 1, 1);
        s = pmSetGpuLaunchConf(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, lc);
    }
    return s;
}

pmStatus d_cpu(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const DConf* c = (const DConf*)ti.taskConf;
    uint32_t* out = (uint32_t*)si.memInfo[0].writePtr;
    for (uint32_t i = 0; i < c->n; ++i) out[i] = (uint32_t)si.subtaskId * c->n + i;
    uint32_t* sc = (uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo,
                                                 SUBTASK_TO_POST_SUBTASK, (size_t)c->orders * 4, NULL);
    if (!sc) return pmMemoryError;
    chunk_lengths(c, (uint32_t)si.subtaskId, sc);
    return pmSuccess;
}

pmStatus d_redist(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const DConf* c = (const DConf*)ti.taskConf;
    const uint32_t* len = (const uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev.deviceHandle, si.subtaskId,
                                                              si.splitInfo, SUBTASK_TO_POST_SUBTASK, 0, NULL);
    if (!len) return pmFatalError;
    size_t off = 0;
    for (uint32_t j = 0; j < c->orders; ++j) {
        pmStatus s = pmRedistributeData(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, off,
                                        (size_t)len[j] * 4, (uint32_t)((si.subtaskId + j) % c->orders));
        if (s != pmSuccess) return s;
        off += (size_t)len[j] * 4;
    }
    return pmSuccess;
}

pmCallbackHandle g_dcb[3];

}  // namespace

// device: 0 CPU stubs, 1 the HIP stub, 2 both. Returns the task's pmStatus; out receives the
// reduced space (n elements of the type).
// sparse: nonzero makes every subtask set only about one element in 16 (the others 0).
extern "C" int rp_reduce_ex(int op, int type, uint32_t subtasks, uint32_t n, int device, uint32_t sparse, void* out) {
    if (op < 0 || op >= MAX_REDUCTION_OP_TYPES || type < 0 || type >= MAX_REDUCTION_DATA_TYPES || device < 0 || device > 2)
        return pmInvalidIndex;
    pmCallbackHandle cbh;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbackHandle& h = g_rcb[op][type][device];
        if (!h) {
            pmCallbacks cb;
            cb.dataDistribution = r_dd;
            if (device != 1) cb.subtask_cpu = r_cpu;
            if (device != 0) cb.subtask_gpu_hip = kernel_of(type);
            cb.dataReduction = pmGetSubtaskReductionCallbackImpl((pmReductionOpType)op, (pmReductionDataType)type);
            char key[64];
            snprintf(key, sizeof(key), "RPROBE_%d_%d_%d", op, type, device);
            if (pmStatus s = pmRegisterCallbacks(key, cb, &h)) return (int)s;
        }
        cbh = h;
    }
    const size_t bytes = (size_t)n * type_size(type);
    pmMemHandle mh = nullptr;
    if (pmCreateMemory(bytes, &mh) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(mh, &p);
    memset(p, 0, bytes);
    RConf c{n, op, type, sparse};
    pmTaskMem mem[1] = {pmTaskMem(mh, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails d(&c, sizeof(c), mem, 1, cbh, subtasks);
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        pmReleaseTask(th);
    }
    pmFetchMemory(mh);
    pmGetRawMemPtr(mh, &p);
    memcpy(out, p, bytes);
    pmReleaseMemory(mh);
    return (int)s;
}

extern "C" int rp_reduce(int op, int type, uint32_t subtasks, uint32_t n, int device, void* out) {
    return rp_reduce_ex(op, type, subtasks, n, device, 0, out);
}

// out: subtasks * n words, the redistributed space; meta: up to orders (order, bytes) pairs;
// *meta_count = entries pmGetRedistributionMetadata returned.
extern "C" int rp_redistribute(uint32_t subtasks, uint32_t n, uint32_t orders, int device, uint32_t* out,
                               uint32_t* meta, uint32_t* meta_count) {
    if (device < 0 || device > 2 || orders == 0) return pmInvalidIndex;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_dcb[device]) {
            pmCallbacks cb;
            cb.dataDistribution = d_dd;
            if (device != 1) cb.subtask_cpu = d_cpu;
            if (device != 0) cb.subtask_gpu_hip = dp_kernel;
            cb.dataRedistribution = d_redist;
            char key[32];
            snprintf(key, sizeof(key), "DPROBE_%d", device);
            if (pmStatus s = pmRegisterCallbacks(key, cb, &g_dcb[device])) return (int)s;
        }
    }
    const size_t bytes = (size_t)subtasks * n * 4;
    pmMemHandle mh = nullptr;
    if (pmCreateMemory(bytes, &mh) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(mh, &p);
    memset(p, 0xff, bytes);
    DConf c{n, orders};
    pmTaskMem mem[1] = {pmTaskMem(mh, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails d(&c, sizeof(c), mem, 1, g_dcb[device], subtasks);
    d.multiAssignEnabled = false;
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    *meta_count = 0;
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        ulong cnt = 0;
        pmRedistributionMetadata* m = pmGetRedistributionMetadata(th, 0, &cnt);
        for (ulong k = 0; k < cnt && k < orders; ++k) {
            meta[2 * k] = m[k].order;
            meta[2 * k + 1] = m[k].count;
        }
        *meta_count = (uint32_t)cnt;
        pmReleaseTask(th);
    }
    pmFetchMemory(mh);
    pmGetRawMemPtr(mh, &p);
    memcpy(out, p, bytes);
    pmReleaseMemory(mh);
    return (int)s;
}
