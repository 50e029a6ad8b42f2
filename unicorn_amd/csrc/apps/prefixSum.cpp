// prefixSum app (reference testSuite/prefixSum): the inclusive prefix sum of an int32 array, in
// three steps (prefixSum.cpp:225-290):
//   1. a task of ceil(n / E) subtasks (E elements each, ELEMS_PER_SUBTASK = 1,000,000 by default)
//      scans each chunk on its own (prefixSumDataDistribution / prefixSum_cpu, 68-93);
//   2. on the host, the last element of every chunk is read back and scanned into the auxiliary
//      array (PopulateAuxMem, 197-221);
//   3. a task of ceil(n / E) - 1 subtasks adds aux[s] to every element of chunk s + 1
//      (elemAddDataDistribution / elemAdd_cpu, 95-131), READ_WRITE with disjoint subscriptions.
// The reference app is written against an older version of its API (pmSubtaskInfo::inputMem,
// subscriptions without a memory index) and does not compile against its current headers; this
// restates it with the current API: input / output / aux are task memory indices 0 / 1 and the
// auxiliary task's spaces are aux (READ_ONLY) and output (READ_WRITE). Arithmetic is int32 with
// wrap-around, as the reference's int. GPU callbacks: uh_scan_i32 and uh_add_scalar_i32 in place
// of prefixSum_cuda and elemAdd_cuda (prefixSumCUDA.cu:68-165).
#include "app_common.h"

namespace prefixSum {

using namespace pm;

struct prefixSumTaskConf {
    unsigned int arrayLen;
    unsigned int elemsPerSubtask;
};

static unsigned int chunk_len(const prefixSumTaskConf* c, unsigned long s) {
    const unsigned long start = s * (unsigned long)c->elemsPerSubtask;
    return c->arrayLen < start + c->elemsPerSubtask ? (unsigned int)(c->arrayLen - start) : c->elemsPerSubtask;
}

// ---- task 1: per-chunk scan

static pmStatus psDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    const pmSubscriptionInfo s((size_t)si.subtaskId * c->elemsPerSubtask * 4, (size_t)chunk_len(c, si.subtaskId) * 4);
    pmStatus st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 0, READ_SUBSCRIPTION, s);
    if (st == pmSuccess) st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 1, WRITE_SUBSCRIPTION, s);
    return st;
}

static pmStatus ps_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    const int32_t* in = (const int32_t*)si.memInfo[0].ptr;
    int32_t* out = (int32_t*)si.memInfo[1].writePtr;
    const unsigned int n = chunk_len(c, si.subtaskId);
    uint32_t acc = 0;
    for (unsigned int i = 0; i < n; ++i) {
        acc += (uint32_t)in[i];
        out[i] = (int32_t)acc;
    }
    return pmSuccess;
}

static pmStatus ps_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    if (uh_scan_i32((const int32_t*)si.memInfo[0].ptr, (int32_t*)si.memInfo[1].writePtr, chunk_len(c, si.subtaskId),
                    stream) != 0) {
        apps::set_error(std::string("uh_scan_i32: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

// ---- task 3: add the scanned chunk totals

static pmStatus eaDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    pmStatus st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 0, READ_SUBSCRIPTION,
                                      pmSubscriptionInfo((size_t)si.subtaskId * 4, 4));
    if (st == pmSuccess)
        st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 1, READ_WRITE_SUBSCRIPTION,
                                 pmSubscriptionInfo((size_t)(si.subtaskId + 1) * c->elemsPerSubtask * 4,
                                                    (size_t)chunk_len(c, si.subtaskId + 1) * 4));
    return st;
}

static pmStatus ea_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    const uint32_t a = (uint32_t)((const int32_t*)si.memInfo[0].ptr)[0];
    int32_t* out = (int32_t*)si.memInfo[1].ptr;
    const unsigned int n = chunk_len(c, si.subtaskId + 1);
    for (unsigned int i = 0; i < n; ++i) out[i] = (int32_t)((uint32_t)out[i] + a);
    return pmSuccess;
}

static pmStatus ea_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const prefixSumTaskConf* c = (const prefixSumTaskConf*)ti.taskConf;
    if (uh_add_scalar_i32((int32_t*)si.memInfo[1].ptr, chunk_len(c, si.subtaskId + 1), (const int32_t*)si.memInfo[0].ptr,
                          stream) != 0) {
        apps::set_error(std::string("uh_add_scalar_i32: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    prefixSumTaskConf conf;
    pmMemHandle in = nullptr, out = nullptr;
    unsigned long subtasks = 0;
};

static apps::CallbackVariants g_cb_scan, g_cb_add;
static std::mutex g_mtx;

}  // namespace prefixSum

using namespace prefixSum;

extern "C" void* ua_ps_create(uint32_t n, uint32_t elems_per_subtask) {
    if (!n || !elems_per_subtask) {
        apps::set_error("ua_ps_create: empty array or chunk");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks a, b;
        a.dataDistribution = psDataDistribution;
        a.subtask_cpu = ps_cpu;
        a.subtask_gpu_custom = ps_gpu;
        b.dataDistribution = eaDataDistribution;
        b.subtask_cpu = ea_cpu;
        b.subtask_gpu_custom = ea_gpu;
        if (apps::register_variants("PREFIXSUM", a, g_cb_scan) != 0 || apps::register_variants("ELEMADD", b, g_cb_add) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->conf.arrayLen = n;
    in->conf.elemsPerSubtask = elems_per_subtask;
    in->subtasks = (n + elems_per_subtask - 1) / elems_per_subtask;
    if (pmCreateMemory((size_t)n * 4, &in->in) != pmSuccess || pmCreateMemory((size_t)n * 4, &in->out) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        if (in->in) pmReleaseMemory(in->in);
        delete in;
        return nullptr;
    }
    return in;
}

extern "C" int ua_ps_set_input(void* h, const int32_t* data) {
    Instance* in = (Instance*)h;
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->in, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, data, (size_t)in->conf.arrayLen * 4);
    return 0;
}

extern "C" int ua_ps_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    const double t0 = apps::now_secs();
    pmTaskMem mem[2] = {pmTaskMem(in->in, READ_ONLY, SUBSCRIPTION_NATURAL), pmTaskMem(in->out, WRITE_ONLY, SUBSCRIPTION_NATURAL)};
    pmTaskDetails d(&in->conf, sizeof(in->conf), mem, 2, g_cb_scan.h[device], in->subtasks);
    d.multiAssignEnabled = apps::split_request().multi_assign;
    int rc = apps::run_task(d, nullptr);
    if (rc || in->subtasks < 2) {
        if (seconds) *seconds = apps::now_secs() - t0;
        return rc;
    }
    // PopulateAuxMem: the last element of every chunk but the last, scanned on the host
    const unsigned long m = in->subtasks - 1;
    pmMemHandle aux = nullptr;
    if (pmCreateMemory(m * 4, &aux) != pmSuccess) return (int)pmMemoryError;
    pmRawMemPtr pa;
    pmGetRawMemPtr(aux, &pa);
    uint32_t acc = 0;
    for (unsigned long s = 0; s < m; ++s) {
        const size_t off = ((s + 1) * (size_t)in->conf.elemsPerSubtask - 1) * 4;
        int32_t v = 0;
        pmStatus st = pmReadMemoryRange(in->out, off, 4, &v);
        if (st != pmSuccess) {
            pmReleaseMemory(aux);
            return (int)st;
        }
        acc += (uint32_t)v;
        ((int32_t*)pa)[s] = (int32_t)acc;
    }
    pmTaskMem mem2[2] = {pmTaskMem(aux, READ_ONLY, SUBSCRIPTION_NATURAL), pmTaskMem(in->out, READ_WRITE, SUBSCRIPTION_NATURAL, true)};
    pmTaskDetails d2(&in->conf, sizeof(in->conf), mem2, 2, g_cb_add.h[device], m);
    d2.multiAssignEnabled = apps::split_request().multi_assign;
    rc = apps::run_task(d2, nullptr);
    pmReleaseMemory(aux);
    if (seconds) *seconds = apps::now_secs() - t0;
    return rc;
}

extern "C" int ua_ps_get_output(void* h, int32_t* out) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->out);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->out, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(out, p, (size_t)in->conf.arrayLen * 4);
    return 0;
}

extern "C" int ua_ps_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    pmReleaseMemory(in->in);
    pmReleaseMemory(in->out);
    delete in;
    return 0;
}
