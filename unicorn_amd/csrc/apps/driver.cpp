// Process-level C-ABI of the app drivers: runtime start/stop, HIP stub controls, counters.
#include "app_common.h"
#include "unicorn_hip.h"

#include <cstdlib>

namespace apps {

static std::mutex g_err_mtx;
static std::string g_err;

void set_error(const std::string& s) {
    std::lock_guard<std::mutex> g(g_err_mtx);
    g_err = s;
}

SplitRequest& split_request() {
    static SplitRequest r;
    return r;
}

}  // namespace apps

using namespace pm;

extern "C" {

int ua_init(int cpu_stubs) {
    if (cpu_stubs > 0) pmSetCpuStubCount(cpu_stubs);
    pmStatus s = pmInitialize();
    if (s != pmSuccess) apps::set_error(pmGetLastError());
    return (int)s;
}

int ua_finalize(void) { return (int)pmFinalize(); }

int ua_has_gpu(void) { return pmHasHipDevice() ? 1 : 0; }

int ua_set_hip_mode(int mode) { return (int)pmSetHipStubMode(mode); }

int ua_set_hip_compact(int compact) {
    return (int)pmSetHipOptimalView(compact ? SUBSCRIPTION_COMPACT : SUBSCRIPTION_NATURAL);
}

int ua_set_hip_graphs(int on) { return (int)pmSetHipGraphs(on != 0); }
int ua_set_hip_depth(int depth) { return (int)pmSetHipPipelineDepth(depth); }
int ua_set_batch_limit(int n) { return (int)pmSetHipBatchLimit(n); }
int ua_set_split(int cpu, int gpu, int cpu_pieces, int gpu_pieces) {
    apps::split_request().cpu = cpu != 0;
    apps::split_request().gpu = gpu != 0;
    return (int)pmSetSplitFactors(cpu_pieces, gpu_pieces);
}
int ua_set_multi_assign(int on) {
    apps::split_request().multi_assign = on != 0;
    return 0;
}
int ua_set_hbm_budget(uint64_t bytes) { return (int)pmSetHbmBudget((size_t)bytes); }
int ua_set_sentinel_threshold(uint64_t bytes, uint64_t* previous) {
    size_t prev = 0;
    const int rc = (int)pmSetSentinelCompressionThreshold((size_t)bytes, &prev);
    if (previous) *previous = prev;
    return rc;
}

int ua_get_stats(ua_stats* s) {
    pmHipStubStats h;
    pmStatus st = pmGetHipStubStats(&h);
    if (st != pmSuccess) return (int)st;
    s->subtasks = h.subtasks;
    s->batches = h.batches;
    s->h2d_bytes = h.h2d_bytes;
    s->d2h_bytes = h.d2h_bytes;
    s->d2d_bytes = h.d2d_bytes;
    s->cache_hits = h.cache_hits;
    s->evictions = h.evictions;
    s->ns_queue = h.ns_queue;
    s->ns_stage = h.ns_stage;
    s->ns_flush = h.ns_flush;
    s->ns_wait = h.ns_wait;
    s->ns_retire = h.ns_retire;
    s->ns_wake = h.ns_wake;
    s->ma_copies = h.ma_copies;
    s->ma_wins = h.ma_wins;
    s->oom_waits = h.oom_waits;
    s->sentinel_compressed = h.sentinel_compressed;
    s->plain_launches = s->persistent_launches = s->queue_items = 0;
    if (pmHasHipDevice()) uh_launch_counters(&s->plain_launches, &s->persistent_launches, &s->queue_items);
    return 0;
}

int ua_reset_stats(void) { return (int)pmResetHipStubStats(); }

int ua_set_kernel_timing(int every) { return every < 0 ? -1 : (int)pmSetHipKernelTimingEvery((unsigned)every); }

int ua_kernel_time(double* ms, uint64_t* flushes, int reset) {
    return (int)pmGetHipKernelTime(ms, flushes, reset != 0);
}

int ua_kernel_time_family(int family, double* ms, uint64_t* flushes, double* work, int reset) {
    return (int)pmGetHipKernelTimeByFamily(family, ms, flushes, work, reset != 0);
}

int ua_device_sync(void) { return pmHasHipDevice() ? uh_device_sync() : 0; }

const char* ua_last_error(void) {
    static thread_local std::string copy;
    {
        std::lock_guard<std::mutex> g(apps::g_err_mtx);
        copy = apps::g_err;
    }
    const char* pe = pmGetLastError();
    if (pe && *pe) copy += (copy.empty() ? "" : " | ") + std::string(pe);
    return copy.c_str();
}

}  // extern "C"
