// Test support (built by unicorn_amd.build.build_test_support into tests/native/librw_probe.so):
// a CPU task over one READ_WRITE space in which subtask s reads block s+1 (cyclically) and
// writes block s, so that one subtask's reads overlap another's writes. Unless the task declares
// its reads and writes disjoint across subtasks, every subtask must read the task's input: the
// writes are held back until the task ends (reference pmTask::DoSubtasksNeedShadowMemory,
// pmTask.cpp:749-754).
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

#include "pmPublicDefinitions.h"
#include "pmExtensions.h"

using namespace pm;

namespace {

struct Conf {
    uint32_t n;          // subtasks
    uint32_t m;          // elements per block
};

pmStatus dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    const size_t b = (size_t)c->m * 4;
    const size_t src = ((si.subtaskId + 1) % c->n) * b, dst = si.subtaskId * b;
    pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, READ_SUBSCRIPTION,
                        pmSubscriptionInfo(src, b));
    pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION,
                        pmSubscriptionInfo(dst, b));
    return pmSuccess;
}

// out block s = in block s+1, plus 1; staggered so that neighbours overlap in time
pmStatus shift_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    const pmMemInfo& mi = si.memInfo[0];
    const uint32_t* src = (const uint32_t*)mi.readPtr;
    uint32_t* dst = (uint32_t*)mi.writePtr;
    if (!src || !dst) return pmFatalError;
    std::this_thread::sleep_for(std::chrono::milliseconds((si.subtaskId % 3) * 2));
    for (uint32_t i = 0; i < c->m; ++i) dst[i] = src[i] + 1;
    return pmSuccess;
}

std::mutex g_mtx;
pmCallbackHandle g_cb = nullptr;

}  // namespace

// data: n*m words, replaced by the task's result. Returns the task's pmStatus.
extern "C" int rw_shift(uint32_t n, uint32_t m, int disjoint, int compact, uint32_t* data) {
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_cb) {
            pmCallbacks cb;
            cb.dataDistribution = dd;
            cb.subtask_cpu = shift_cpu;
            if (pmStatus s = pmRegisterCallbacks("RWPROBE", cb, &g_cb)) return (int)s;
        }
    }
    const size_t bytes = (size_t)n * m * 4;
    pmMemHandle h = nullptr;
    if (pmCreateMemory(bytes, &h) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(h, &p);
    memcpy(p, data, bytes);
    Conf c{n, m};
    pmTaskMem mem[1] = {pmTaskMem(h, READ_WRITE, compact ? SUBSCRIPTION_COMPACT : SUBSCRIPTION_NATURAL, disjoint != 0)};
    pmTaskDetails d(&c, sizeof(c), mem, 1, g_cb, n);
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        pmReleaseTask(th);
    }
    pmFetchMemory(h);
    pmGetRawMemPtr(h, &p);
    memcpy(data, p, bytes);
    pmReleaseMemory(h);
    return (int)s;
}
