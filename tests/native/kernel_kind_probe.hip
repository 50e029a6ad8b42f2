// Test support (built by unicorn_amd.build.build_test_support into tests/native/libkkprobe.so):
// a task whose GPU callback is of the kernel kind (subtask_gpu_hip), exercising what the
// reference's kernel-kind contract hands to a __global__ callback (pmCudaInterface.cu:532-543,
// pmExecutionStub.cpp:3300-3368, pmDispatcherGPU.cpp:161-205):
//   - taskInfo.taskConf is a device copy of the task configuration;
//   - pDeviceInfo is a device pointer to the stub's pmDeviceInfo;
//   - the launch geometry is the one set with pmSetGpuLaunchConf;
//   - gpuContext.reservedGlobalMem is device memory of the size pmReserveGpuGlobalMem asked for;
//   - pmGetScratchBuffer in device code returns the PRE_SUBTASK_TO_SUBTASK buffer that the
//     data-distribution callback filled on the host;
//   - memInfo[0].writePtr is the subtask's slice of the output.
// The CPU callback computes the same values from the same scratch buffer (taken inside
// subtask_cpu), without the GPU-only terms.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "pmPublicDefinitions.h"

using namespace pm;

namespace {

struct ProbeConf {
    uint32_t n;          // elements per subtask
    uint32_t magic;
    uint32_t threads;    // threads per block of the launch
    uint32_t pad;
};

pmStatus probe_dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const ProbeConf* c = (const ProbeConf*)ti.taskConf;
    pmSubscriptionInfo w;
    w.offset = (size_t)si.subtaskId * c->n * 4;
    w.length = (size_t)c->n * 4;
    pmStatus s = pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION, w);
    if (s != pmSuccess) return s;
    uint32_t* sc = (uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo,
                                                 PRE_SUBTASK_TO_SUBTASK, (size_t)c->n * 4, NULL);
    if (!sc) return pmMemoryError;
    for (uint32_t i = 0; i < c->n; ++i) sc[i] = (uint32_t)si.subtaskId * 1000003u + i;
    if (dev.deviceType == GPU_HIP) {
        pmGpuLaunchConf lc((int)((c->n + c->threads - 1) / c->threads), 1, 1, (int)c->threads, 1, 1);
        s = pmSetGpuLaunchConf(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, lc);
        if (s == pmSuccess) s = pmReserveGpuGlobalMem(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, (size_t)c->n * 4);
    }
    return s;
}

__global__ void probe_kernel(pmTaskInfo ti, pmDeviceInfo* dev, pmSubtaskInfo si, pmStatus* status) {
    const ProbeConf* c = (const ProbeConf*)ti.taskConf;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c->n) return;
    const uint32_t* sc = (const uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev->deviceHandle, si.subtaskId, si.splitInfo,
                                                             PRE_SUBTASK_TO_SUBTASK, 0, &si.gpuContext);
    uint32_t* res = (uint32_t*)si.gpuContext.reservedGlobalMem;
    res[i] = sc[i] ^ c->magic;
    uint32_t* out = (uint32_t*)si.memInfo[0].writePtr;
    out[i] = res[i] + (dev->deviceType == GPU_HIP ? 1u : 0u) + blockDim.x * 16u + (uint32_t)ti.subtaskCount;
    if (i == 0) *status = pmSuccess;
}

pmStatus probe_cpu(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const ProbeConf* c = (const ProbeConf*)ti.taskConf;
    const uint32_t* sc = (const uint32_t*)pmGetScratchBuffer(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo,
                                                             PRE_SUBTASK_TO_SUBTASK, 0, NULL);
    if (!sc) return pmMemoryError;
    uint32_t* out = (uint32_t*)si.memInfo[0].writePtr;
    for (uint32_t i = 0; i < c->n; ++i) out[i] = (sc[i] ^ c->magic) + c->threads * 16u + (uint32_t)ti.subtaskCount;
    return pmSuccess;
}

std::mutex g_mtx;
pmCallbackHandle g_cb[2] = {nullptr, nullptr};   // CPU, GPU

}  // namespace

// Runs one probe task of `subtasks` subtasks of n elements on the CPU stubs (device 0) or the
// HIP stub (device 1) and copies the output (n * subtasks uint32) to out. Returns the pmStatus.
extern "C" int kk_probe_run(uint32_t n, uint32_t subtasks, uint32_t magic, uint32_t threads, int device, uint32_t* out) {
    if (!n || !subtasks || !threads || threads > 1024 || (device != 0 && device != 1) || !out) return pmInvalidIndex;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_cb[0]) {
            pmCallbacks cpu, gpu;
            cpu.dataDistribution = gpu.dataDistribution = probe_dd;
            cpu.subtask_cpu = probe_cpu;
            gpu.subtask_gpu_hip = probe_kernel;
            pmStatus s = pmRegisterCallbacks("KKPROBE_CPU", cpu, &g_cb[0]);
            if (s == pmSuccess) s = pmRegisterCallbacks("KKPROBE_GPU", gpu, &g_cb[1]);
            if (s != pmSuccess) return s;
        }
    }
    const size_t bytes = (size_t)n * subtasks * 4;
    pmMemHandle mem = nullptr;
    pmStatus s = pmCreateMemory(bytes, &mem);
    if (s != pmSuccess) return s;
    ProbeConf conf{n, magic, threads, 0};
    pmTaskMem tm[1] = {pmTaskMem(mem, WRITE_ONLY, SUBSCRIPTION_OPTIMAL)};
    pmTaskDetails d(&conf, sizeof(conf), tm, 1, g_cb[device], subtasks);
    pmTaskHandle th = nullptr;
    s = pmSubmitTask(d, &th);
    if (s == pmSuccess) s = pmWaitForTaskCompletion(th);
    if (th) pmReleaseTask(th);
    if (s == pmSuccess) s = pmFetchMemory(mem);
    pmRawMemPtr p = nullptr;
    if (s == pmSuccess) s = pmGetRawMemPtr(mem, &p);
    if (s == pmSuccess) memcpy(out, p, bytes);
    pmReleaseMemory(mem);
    return s;
}
