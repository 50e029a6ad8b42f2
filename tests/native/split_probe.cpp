// Test support (built by unicorn_amd.build.build_test_support into tests/native/libsplit_probe.so):
// a CPU task submitted with canSplitCpuSubtasks, whose pieces record the (subtask, splitId,
// splitCount) they were called with and fill their share of the subtask's output rows (the
// reference's split subtasks, pmSubtaskSplitter.cpp:49-140, with the apps' row split of
// matrixMultiplyBlas.cpp:67-88).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "pmPublicDefinitions.h"
#include "pmExtensions.h"

using namespace pm;

namespace {

struct Conf {
    uint32_t rows;       // rows per subtask (one uint32 each)
    uint32_t sleep_ms;   // each piece sleeps this long first
};

// Rows [first, first + count) of a subtask handled by one piece; false for a piece with no rows.
bool piece_rows(const Conf* c, const pmSplitInfo& s, uint32_t* first, uint32_t* count) {
    *first = 0;
    *count = c->rows;
    if (!s.splitCount) return true;
    const uint32_t n = std::min(c->rows, s.splitCount);
    if (s.splitId >= n) return false;
    const uint32_t part = c->rows / n;
    *first = s.splitId * part;
    *count = s.splitId == n - 1 ? c->rows - *first : part;
    return true;
}

std::mutex g_mtx;
std::vector<uint32_t> g_log;   // subtask, splitId, splitCount per call

pmStatus dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    uint32_t f, n;
    if (!piece_rows(c, si.splitInfo, &f, &n)) return pmSuccess;
    pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0, WRITE_SUBSCRIPTION,
                        pmSubscriptionInfo(((size_t)si.subtaskId * c->rows + f) * 4, (size_t)n * 4));
    return pmSuccess;
}

pmStatus piece_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        g_log.push_back((uint32_t)si.subtaskId);
        g_log.push_back(si.splitInfo.splitId);
        g_log.push_back(si.splitInfo.splitCount);
    }
    if (c->sleep_ms) std::this_thread::sleep_for(std::chrono::milliseconds(c->sleep_ms));
    uint32_t f, n;
    if (!piece_rows(c, si.splitInfo, &f, &n)) return pmSuccess;
    uint32_t* out = (uint32_t*)si.memInfo[0].ptr;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = ((uint32_t)si.subtaskId << 16) | (f + i);
    return pmSuccess;
}

pmCallbackHandle g_cb = nullptr;

}  // namespace

// Runs one task of `subtasks` subtasks x `rows` rows with CPU splitting into `pieces` per subtask
// (0: one per CPU stub). out: rows*subtasks words (subtask << 16 | row in subtask, or 0 where no
// piece wrote). log: up to log_cap words, 3 per callback call; *log_len = words recorded.
// Returns the task's pmStatus.
extern "C" int split_run(uint32_t subtasks, uint32_t rows, uint32_t pieces, uint32_t sleep_ms, int timeout_secs,
                         uint32_t* out, uint32_t* log, uint32_t log_cap, uint32_t* log_len) {
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_cb) {
            pmCallbacks cb;
            cb.dataDistribution = dd;
            cb.subtask_cpu = piece_cpu;
            if (pmStatus s = pmRegisterCallbacks("SPLITPROBE", cb, &g_cb)) return (int)s;
        }
        g_log.clear();
    }
    if (pmStatus s = pmSetSplitFactors((int)pieces, 0)) return (int)s;
    const size_t bytes = (size_t)subtasks * rows * 4;
    pmMemHandle h = nullptr;
    if (pmCreateMemory(bytes, &h) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(h, &p);
    memset(p, 0, bytes);
    Conf c{rows, sleep_ms};
    pmTaskMem m[1] = {pmTaskMem(h, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails d(&c, sizeof(c), m, 1, g_cb, subtasks);
    d.canSplitCpuSubtasks = true;
    if (timeout_secs >= 0) d.timeOutInSecs = timeout_secs;
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        pmReleaseTask(th);
    }
    pmSetSplitFactors(0, 0);
    pmFetchMemory(h);
    pmGetRawMemPtr(h, &p);
    memcpy(out, p, bytes);
    pmReleaseMemory(h);
    std::lock_guard<std::mutex> g(g_mtx);
    *log_len = (uint32_t)g_log.size();
    memcpy(log, g_log.data(), std::min<size_t>(g_log.size(), log_cap) * 4);
    return (int)s;
}
