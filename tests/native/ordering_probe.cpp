// Test support (built by unicorn_amd.build.build_test_support into tests/native/libordering_probe.so):
// CPU-only tasks that share address spaces, submitted back to back without waiting, to observe
// the runtime's ordering of tasks on one address space (reference: address-space locks,
// pmAddressSpace.cpp:500-530, pmTask.cpp:176-193).
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

#include "pmPublicDefinitions.h"
#include "pmExtensions.h"

using namespace pm;

namespace {

struct Conf {
    uint32_t n;          // elements per subtask
    uint32_t value;      // what the writer stores
    uint32_t sleep_ms;   // each subtask sleeps this long first
    uint32_t pad;
};

double now() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

std::atomic<int> g_running{0}, g_max_running{0};
std::mutex g_tmtx;
double g_writer_end = 0, g_reader_start = 1e300;

pmStatus dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    pmSubscriptionInfo r((size_t)si.subtaskId * c->n * 4, (size_t)c->n * 4);
    for (uint i = 0; i < si.memCount; ++i)
        pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, i,
                            i == 0 && si.memCount == 2 ? READ_SUBSCRIPTION : WRITE_SUBSCRIPTION, r);
    return pmSuccess;
}

// writer: one space (WRITE_ONLY), filled with value after a sleep
pmStatus writer_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    std::this_thread::sleep_for(std::chrono::milliseconds(c->sleep_ms));
    uint32_t* out = (uint32_t*)si.memInfo[0].ptr;
    for (uint32_t i = 0; i < c->n; ++i) out[i] = c->value;
    std::lock_guard<std::mutex> g(g_tmtx);
    g_writer_end = std::max(g_writer_end, now());
    return pmSuccess;
}

// reader: space 0 (READ_ONLY) copied into space 1 (WRITE_ONLY); counts concurrent readers
pmStatus reader_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    {
        std::lock_guard<std::mutex> g(g_tmtx);
        g_reader_start = std::min(g_reader_start, now());
    }
    const int r = ++g_running;
    int m = g_max_running.load();
    while (r > m && !g_max_running.compare_exchange_weak(m, r)) {}
    std::this_thread::sleep_for(std::chrono::milliseconds(c->sleep_ms));
    memcpy(si.memInfo[1].ptr, si.memInfo[0].ptr, (size_t)c->n * 4);
    --g_running;
    return pmSuccess;
}

pmCallbackHandle g_w = nullptr, g_r = nullptr;
std::mutex g_mtx;

int ensure_callbacks() {
    std::lock_guard<std::mutex> g(g_mtx);
    if (g_w) return 0;
    pmCallbacks w, r;
    w.dataDistribution = r.dataDistribution = dd;
    w.subtask_cpu = writer_cpu;
    r.subtask_cpu = reader_cpu;
    pmStatus s = pmRegisterCallbacks("ORDPROBE_W", w, &g_w);
    if (s == pmSuccess) s = pmRegisterCallbacks("ORDPROBE_R", r, &g_r);
    return (int)s;
}

}  // namespace

// A writer of `space` (sleep_ms per subtask) and then a reader of it, submitted back to back
// without waiting (the reader unordered if asked). out: what the reader copied. gap: reader start
// minus writer end, in seconds (negative when the reader started first).
extern "C" int ord_writer_reader(int unordered, uint32_t n, uint32_t subtasks, uint32_t value, uint32_t sleep_ms,
                                 uint32_t* out, double* gap) {
    if (int rc = ensure_callbacks()) return rc;
    g_writer_end = 0;
    g_reader_start = 1e300;
    pmMemHandle a = nullptr, b = nullptr;
    const size_t bytes = (size_t)n * subtasks * 4;
    if (pmCreateMemory(bytes, &a) != pmSuccess || pmCreateMemory(bytes, &b) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(a, &p);
    memset(p, 0, bytes);
    Conf cw{n, value, sleep_ms, 0}, cr{n, 0, 1, 0};
    pmTaskMem mw[1] = {pmTaskMem(a, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskMem mr[2] = {pmTaskMem(a, READ_ONLY, SUBSCRIPTION_NATURAL), pmTaskMem(b, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails dw(&cw, sizeof(cw), mw, 1, g_w, subtasks), dr(&cr, sizeof(cr), mr, 2, g_r, subtasks);
    pmTaskHandle hw = nullptr, hr = nullptr;
    pmTaskSubmitOptions opt{-1, unordered != 0};
    pmStatus s = pmSubmitTask(dw, &hw);
    if (s == pmSuccess) s = pmSubmitTaskWithOptions(dr, &opt, &hr);
    pmStatus s1 = hw ? pmWaitForTaskCompletion(hw) : pmSuccess;
    pmStatus s2 = hr ? pmWaitForTaskCompletion(hr) : pmSuccess;
    if (hw) pmReleaseTask(hw);
    if (hr) pmReleaseTask(hr);
    if (s == pmSuccess) s = s1 != pmSuccess ? s1 : s2;
    if (s == pmSuccess) s = pmFetchMemory(b);
    if (s == pmSuccess) s = pmGetRawMemPtr(b, &p);
    if (s == pmSuccess) memcpy(out, p, bytes);
    if (gap) *gap = g_reader_start - g_writer_end;
    pmReleaseMemory(a);
    pmReleaseMemory(b);
    return (int)s;
}

// `readers` reader tasks of one shared space, submitted back to back: returns the largest number
// of reader subtasks that ran at the same time.
extern "C" int ord_readers_overlap(uint32_t readers, uint32_t sleep_ms, int* max_running) {
    if (int rc = ensure_callbacks()) return rc;
    g_running = 0;
    g_max_running = 0;
    const uint32_t n = 1024;
    pmMemHandle a = nullptr;
    if (pmCreateMemory((size_t)n * 4, &a) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(a, &p);
    memset(p, 0, (size_t)n * 4);
    Conf cr{n, 0, sleep_ms, 0};
    pmMemHandle outs[16] = {};
    pmTaskHandle hs[16] = {};
    pmStatus s = pmSuccess;
    readers = readers > 16 ? 16 : readers;
    for (uint32_t i = 0; i < readers && s == pmSuccess; ++i) {
        s = pmCreateMemory((size_t)n * 4, &outs[i]);
        if (s != pmSuccess) break;
        pmTaskMem mr[2] = {pmTaskMem(a, READ_ONLY, SUBSCRIPTION_NATURAL), pmTaskMem(outs[i], WRITE_ONLY, SUBSCRIPTION_NATURAL)};
        pmTaskDetails dr(&cr, sizeof(cr), mr, 2, g_r, 1);
        s = pmSubmitTask(dr, &hs[i]);
    }
    for (uint32_t i = 0; i < readers; ++i) {
        if (hs[i]) {
            pmStatus w = pmWaitForTaskCompletion(hs[i]);
            if (s == pmSuccess) s = w;
            pmReleaseTask(hs[i]);
        }
        if (outs[i]) pmReleaseMemory(outs[i]);
    }
    pmReleaseMemory(a);
    *max_running = g_max_running.load();
    return (int)s;
}
