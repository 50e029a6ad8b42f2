// Test support (built by unicorn_amd.build.build_test_support into tests/native/libma_probe.so):
// a task that both stub kinds run (multi-assign), with a slow CPU callback (it sleeps first) and
// a fast GPU callback (one uh_fill_u32 over the subtask's block), so that the idle HIP stub
// copies the subtasks the CPU stubs still hold (reference pmExecutionStub.cpp:1058-1170). Every
// subtask s fills its block with s * 7 + 1 (+ base), whichever copy commits.
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

#include "pmPublicDefinitions.h"
#include "pmExtensions.h"
#include "unicorn_hip.h"

using namespace pm;

namespace {

struct Conf {
    uint32_t n;          // words per subtask block
    uint32_t sleep_ms;   // CPU callback delay
    uint32_t base;
    uint32_t pad;
};

pmStatus dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    return pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION,
                               pmSubscriptionInfo((size_t)si.subtaskId * c->n * 4, (size_t)c->n * 4));
}

pmStatus slow_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    std::this_thread::sleep_for(std::chrono::milliseconds(c->sleep_ms));
    uint32_t* out = (uint32_t*)si.memInfo[0].writePtr;
    for (uint32_t i = 0; i < c->n; ++i) out[i] = (uint32_t)si.subtaskId * 7u + 1u + c->base;
    return pmSuccess;
}

pmStatus fast_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const Conf* c = (const Conf*)ti.taskConf;
    return uh_fill_u32(si.memInfo[0].writePtr, (uint32_t)si.subtaskId * 7u + 1u + c->base, c->n, stream) ? pmGraphicsCardError
                                                                                                          : pmSuccess;
}

std::mutex g_mtx;
pmCallbackHandle g_cb = nullptr;

}  // namespace

// Returns the task's pmStatus; out: subtasks * n words; stats: the HIP stub's counters after it.
extern "C" int ma_run(uint32_t subtasks, uint32_t n, uint32_t sleep_ms, uint32_t base, int multi_assign,
                      uint32_t* out, uint64_t* copies, uint64_t* wins) {
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_cb) {
            pmCallbacks cb;
            cb.dataDistribution = dd;
            cb.subtask_cpu = slow_cpu;
            cb.subtask_gpu_custom = fast_gpu;
            if (pmStatus s = pmRegisterCallbacks("MAPROBE", cb, &g_cb)) return (int)s;
        }
    }
    const size_t bytes = (size_t)subtasks * n * 4;
    pmMemHandle h = nullptr;
    if (pmCreateMemory(bytes, &h) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(h, &p);
    memset(p, 0, bytes);
    Conf c{n, sleep_ms, base, 0};
    pmTaskMem mem[1] = {pmTaskMem(h, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails d(&c, sizeof(c), mem, 1, g_cb, subtasks);
    d.multiAssignEnabled = multi_assign != 0;
    pmResetHipStubStats();
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        pmReleaseTask(th);
    }
    pmHipStubStats st;
    if (pmGetHipStubStats(&st) == pmSuccess) {
        *copies = st.ma_copies;
        *wins = st.ma_wins;
    }
    pmFetchMemory(h);
    pmGetRawMemPtr(h, &p);
    memcpy(out, p, bytes);
    pmReleaseMemory(h);
    return (int)s;
}
