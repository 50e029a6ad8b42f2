// pageRank in the reference's own decomposition (testSuite/pageRank/source/code/src/pageRank.cpp:
// 237-378, 449-560, pageRankCUDA.cu:46-164): subtasks own *source* pages, read their padded
// out-link records, and push 0.85 * x / outdeg into a private full-length output vector; the
// built-in REDUCE_ADD over floats folds the subtasks' vectors (pmReducer.cpp:593-605). The pull
// app (pageRank.cpp) is the fast path; this one is the drop-in for a caller that keeps the
// reference's callbacks, and it exercises the runtime's reductions: on the HIP stub's resident
// modes the per-subtask vectors are HBM shadows folded with uh_reduce_add_f32, in PLAIN mode they
// are folded on the host after their D2H.
//
// Address spaces of one iteration task:
//   0 records (READ_ONLY): pages x (1 + max_out) uint32, the web dump's records (count, 1-based
//     targets, zero padding), resident in HBM instead of streamed from mapped files each iteration
//   1 x (READ_ONLY): the previous iteration's ranks (not subscribed on iteration 0, which uses
//     x0 = 1.0 like the reference's initialPageRank)
//   2 y (WRITE_ONLY, REDUCE_ADD): this iteration's ranks
#include "app_common.h"
#include "unicorn_hip.h"
#include "unicorn_apps.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace pageRankPush {

using namespace pm;

enum memIndex { RECORDS = 0, X_IN = 1, Y_OUT = 2 };

struct PushConf {
    uint32_t pages, max_out, per_subtask, iteration;
    float x0, damping;
};

static void range_of(const PushConf* c, ulong id, uint32_t* s0, uint32_t* s1) {
    *s0 = (uint32_t)(id * c->per_subtask);
    const uint64_t e = (uint64_t)*s0 + c->per_subtask;
    *s1 = (uint32_t)(e < c->pages ? e : c->pages);
}

static pmStatus distribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const PushConf* c = (const PushConf*)ti.taskConf;
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    const size_t rec = ((size_t)c->max_out + 1) * 4;
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, RECORDS, READ_SUBSCRIPTION,
                        pmSubscriptionInfo((size_t)s0 * rec, (size_t)(s1 - s0) * rec));
    if (c->iteration > 0)
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, X_IN, READ_SUBSCRIPTION,
                            pmSubscriptionInfo((size_t)s0 * 4, (size_t)(s1 - s0) * 4));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, Y_OUT, WRITE_SUBSCRIPTION,
                        pmSubscriptionInfo(0, (size_t)c->pages * 4));
    return pmSuccess;
}

// pageRank_cpu (pageRank.cpp:336-378): zero the private vector, push every page's share.
static pmStatus push_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const PushConf* c = (const PushConf*)ti.taskConf;
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    const uint32_t* rec = (const uint32_t*)si.memInfo[RECORDS].ptr;
    const float* x = c->iteration > 0 ? (const float*)si.memInfo[X_IN].ptr : nullptr;
    float* y = (float*)si.memInfo[Y_OUT].ptr;
    memset(y, 0, (size_t)c->pages * 4);
    const size_t stride = (size_t)c->max_out + 1;
    for (uint32_t p = 0; p < s1 - s0; ++p) {
        const uint32_t* r = rec + p * stride;
        const uint32_t n = r[0];
        if (!n) continue;
        const float incr = (c->damping * (x ? x[p] : c->x0)) / (float)n;
        for (uint32_t k = 0; k < n; ++k) y[r[1 + k] - 1] += incr;
    }
    return pmSuccess;
}

// pageRank_cudaLaunchFunc (pageRankCUDA.cu:119-164): zeroInit, then the push kernel.
static pmStatus push_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const PushConf* c = (const PushConf*)ti.taskConf;
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    const float* x = c->iteration > 0 ? (const float*)si.memInfo[X_IN].ptr : nullptr;
    float* y = (float*)si.memInfo[Y_OUT].ptr;
    int rc = uh_fill_u32(y, 0u, c->pages, stream);
    if (!rc)
        rc = uh_pagerank_push_ell((const uint32_t*)si.memInfo[RECORDS].ptr, s1 - s0, c->max_out, x, c->x0,
                                  c->iteration == 0, c->damping, y, stream);
    if (rc) {
        apps::set_error(std::string("pageRank push GPU callback: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    uint32_t pages = 0, max_out = 0, per_subtask = 0;
    float damping = 0.85f;
    pmMemHandle records = nullptr, state[2] = {nullptr, nullptr};
    int last = -1;                  // state holding the latest ranks
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

static Instance* create(uint32_t pages, uint32_t max_out, uint32_t per_subtask, float damping) {
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = distribution;
        cb.subtask_cpu = push_cpu;
        cb.subtask_gpu_custom = push_gpu;
        cb.dataReduction = pmGetSubtaskReductionCallbackImpl(REDUCE_ADD, REDUCE_FLOATS);
        if (apps::register_variants("PAGERANK_PUSH", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->pages = pages;
    in->max_out = max_out;
    in->per_subtask = per_subtask;
    in->damping = damping;
    const size_t rec_bytes = (size_t)pages * ((size_t)max_out + 1) * 4;
    if (pmCreateMemory(rec_bytes, &in->records) != pmSuccess || pmCreateMemory((size_t)pages * 4, &in->state[0]) != pmSuccess ||
        pmCreateMemory((size_t)pages * 4, &in->state[1]) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        ua_prp_destroy(in);
        return nullptr;
    }
    return in;
}

}  // namespace pageRankPush

using namespace pageRankPush;

// From an out-link CSR (0-based targets); the records are padded to the largest out-degree.
extern "C" void* ua_prp_create(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                               uint32_t pages_per_subtask, float damping) {
    if (!pages || !out_row_ptr || !pages_per_subtask) {
        apps::set_error("ua_prp_create: invalid arguments");
        return nullptr;
    }
    uint32_t max_out = 1;
    for (uint32_t p = 0; p < pages; ++p) max_out = std::max(max_out, out_row_ptr[p + 1] - out_row_ptr[p]);
    Instance* in = create(pages, max_out, pages_per_subtask, damping);
    if (!in) return nullptr;
    pmRawMemPtr raw;
    pmGetRawMemPtr(in->records, &raw);
    uint32_t* rec = (uint32_t*)raw;
    const size_t stride = (size_t)max_out + 1;
    memset(rec, 0, (size_t)pages * stride * 4);
    for (uint32_t p = 0; p < pages; ++p) {
        uint32_t* r = rec + p * stride;
        r[0] = out_row_ptr[p + 1] - out_row_ptr[p];
        for (uint32_t k = out_row_ptr[p]; k < out_row_ptr[p + 1]; ++k) r[1 + k - out_row_ptr[p]] = out_col[k] + 1;
    }
    return in;
}

// From a web dump: the records address space is the dump's page files, copied as they are from
// their mappings (in parallel over files); nothing is converted.
extern "C" void* ua_prp_create_from_webdump(const char* dir, uint32_t pages_per_subtask, float damping) {
    uint32_t pages, max_out, per_file;
    if (ua_webdump_meta(dir, &pages, &max_out, &per_file)) return nullptr;
    if (!pages_per_subtask) {
        apps::set_error("ua_prp_create_from_webdump: zero pages per subtask");
        return nullptr;
    }
    Instance* in = create(pages, max_out, pages_per_subtask, damping);
    if (!in) return nullptr;
    pmRawMemPtr raw;
    pmGetRawMemPtr(in->records, &raw);
    uint8_t* dst = (uint8_t*)raw;
    const size_t rec = ((size_t)max_out + 1) * 4;
    const uint64_t files = (pages + (uint64_t)per_file - 1) / per_file;
    std::atomic<uint64_t> next{0};
    std::atomic<int> bad{0};
    std::vector<std::thread> th;
    const unsigned nt = (unsigned)std::min<uint64_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())), files);
    std::string err;
    std::mutex err_mtx;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (uint64_t f; !bad.load() && (f = next.fetch_add(1)) < files;) {
                const uint64_t p0 = f * per_file, n = std::min<uint64_t>(per_file, pages - p0);
                const std::string path = std::string(dir) + "/web/page_" + std::to_string(p0 + 1);
                const void* src = nullptr;
                if (pmMapFile(path.c_str()) == pmSuccess) src = pmGetMappedFile(path.c_str());
                if (!src) {
                    std::lock_guard<std::mutex> g(err_mtx);
                    err = "web dump: cannot map " + path;
                    bad.store(1);
                    break;
                }
                memcpy(dst + p0 * rec, src, n * rec);
                pmUnmapFile(path.c_str());
            }
        });
    for (auto& x : th) x.join();
    if (bad.load()) {
        apps::set_error(err);
        ua_prp_destroy(in);
        return nullptr;
    }
    // the records are trusted only after the same checks as ua_webdump_read
    const uint32_t* r = (const uint32_t*)raw;
    for (uint64_t p = 0; p < pages; ++p) {
        const uint32_t* q = r + p * (rec / 4);
        bool ok = q[0] <= max_out;
        for (uint32_t k = 0; ok && k < q[0]; ++k) ok = q[1 + k] >= 1 && q[1 + k] <= pages;
        if (!ok) {
            apps::set_error("web dump: invalid record of page " + std::to_string(p + 1));
            ua_prp_destroy(in);
            return nullptr;
        }
    }
    return in;
}

// Runs `iterations` iteration tasks (iteration 0 from x0 = 1.0) on the device set.
extern "C" int ua_prp_run(void* h, int iterations, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (!in || device < 0 || device > 2 || iterations < 1) return -1;
    double t = 0, s = 0;
    int rc = 0;
    for (int it = 0; it < iterations && rc == 0; ++it) {
        PushConf conf{in->pages, in->max_out, in->per_subtask, (uint32_t)it, 1.0f, in->damping};
        const int out = it % 2, inp = 1 - out;
        pmTaskMem mem[3];
        mem[RECORDS] = pmTaskMem(in->records, READ_ONLY, SUBSCRIPTION_OPTIMAL);
        mem[X_IN] = pmTaskMem(in->state[inp], READ_ONLY, SUBSCRIPTION_OPTIMAL);
        mem[Y_OUT] = pmTaskMem(in->state[out], WRITE_ONLY, SUBSCRIPTION_NATURAL);
        const ulong subtasks = (in->pages + in->per_subtask - 1) / in->per_subtask;
        pmTaskDetails d(&conf, sizeof(conf), mem, 3, g_cb.h[device], subtasks);
        rc = apps::run_task(d, &s);
        t += s;
        if (rc == 0) in->last = out;
    }
    if (seconds) *seconds = t;
    return rc;
}

extern "C" int ua_prp_get_ranks(void* h, float* y) {
    Instance* in = (Instance*)h;
    if (!in || in->last < 0) return -1;
    pmMemHandle m = in->state[in->last];
    pmStatus st = pmFetchMemory(m);
    if (st != pmSuccess) return (int)st;
    pmRawMemPtr p;
    st = pmGetRawMemPtr(m, &p);
    if (st != pmSuccess) return (int)st;
    memcpy(y, p, (size_t)in->pages * 4);
    return 0;
}

extern "C" int ua_prp_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    if (in->records) pmReleaseMemory(in->records);
    if (in->state[0]) pmReleaseMemory(in->state[0]);
    if (in->state[1]) pmReleaseMemory(in->state[1]);
    delete in;
    return 0;
}
