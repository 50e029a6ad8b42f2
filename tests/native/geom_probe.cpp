// Test support (built by unicorn_amd.build.build_test_support into tests/native/libgeom_probe.so):
// one CPU subtask over one READ_WRITE space of L bytes whose byte i holds (i * 131 + 7) mod 251,
// subscribed with a list of contiguous and scattered READ / WRITE / READ_WRITE subscriptions.
// The callback reports what its view holds: length, readPtr and writePtr offsets from ptr, and
// the bytes of the view; then it fills the whole view with 0xEE, so that the bytes the runtime
// writes back (the write records only) show in the space after the task.
#include <cstring>
#include <mutex>
#include <vector>

#include "pmPublicDefinitions.h"
#include "pmExtensions.h"

using namespace pm;

namespace {

struct Sub {
    uint64_t type, offset, size, step, count;    // step 0: contiguous (offset, size)
};

struct Conf {
    uint32_t nsubs;
    uint32_t pad;
    Sub subs[32];
};

struct Report {
    uint64_t length, read_off, write_off;
    int64_t has_read, has_write;
    std::vector<uint8_t> bytes;
};
Report g_rep;
std::mutex g_mtx;

pmStatus dd(pmTaskInfo ti, pmDeviceInfo dev, pmSubtaskInfo si) {
    const Conf* c = (const Conf*)ti.taskConf;
    for (uint32_t k = 0; k < c->nsubs; ++k) {
        const Sub& s = c->subs[k];
        pmStatus st = s.step ? pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0,
                                                   (pmSubscriptionType)s.type,
                                                   pmScatteredSubscriptionInfo(s.offset, s.size, s.step, s.count))
                             : pmSubscribeToMemory(ti.taskHandle, dev.deviceHandle, si.subtaskId, si.splitInfo, 0,
                                                   (pmSubscriptionType)s.type, pmSubscriptionInfo(s.offset, s.size));
        if (st != pmSuccess) return st;
    }
    return pmSuccess;
}

pmStatus cpu(pmTaskInfo, pmDeviceInfo, pmSubtaskInfo si) {
    const pmMemInfo& m = si.memInfo[0];
    std::lock_guard<std::mutex> g(g_mtx);
    g_rep.length = m.length;
    g_rep.has_read = m.readPtr != nullptr;
    g_rep.has_write = m.writePtr != nullptr;
    g_rep.read_off = m.readPtr ? (uint64_t)((uint8_t*)m.readPtr - (uint8_t*)m.ptr) : 0;
    g_rep.write_off = m.writePtr ? (uint64_t)((uint8_t*)m.writePtr - (uint8_t*)m.ptr) : 0;
    g_rep.bytes.assign((uint8_t*)m.ptr, (uint8_t*)m.ptr + m.length);
    memset(m.ptr, 0xEE, m.length);
    return pmSuccess;
}

pmCallbackHandle g_cb = nullptr;

}  // namespace

// subs: nsubs x 5 uint64 (type, offset, size, step, count). info: length, read_off, write_off,
// has_read, has_write. view: the view's bytes (up to view_cap). space: the L bytes after the task.
extern "C" int geom_run(uint64_t L, int compact, const uint64_t* subs, uint32_t nsubs, uint64_t* info, uint8_t* view,
                        uint64_t view_cap, uint8_t* space) {
    if (nsubs > 32) return pmInvalidIndex;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        if (!g_cb) {
            pmCallbacks cb;
            cb.dataDistribution = dd;
            cb.subtask_cpu = cpu;
            if (pmStatus s = pmRegisterCallbacks("GEOMPROBE", cb, &g_cb)) return (int)s;
        }
    }
    pmMemHandle h = nullptr;
    if (pmCreateMemory(L, &h) != pmSuccess) return pmMemoryError;
    pmRawMemPtr p;
    pmGetRawMemPtr(h, &p);
    for (uint64_t i = 0; i < L; ++i) ((uint8_t*)p)[i] = (uint8_t)((i * 131 + 7) % 251);
    Conf c;
    memset(&c, 0, sizeof(c));
    c.nsubs = nsubs;
    memcpy(c.subs, subs, nsubs * sizeof(Sub));
    pmTaskMem mem[1] = {pmTaskMem(h, READ_WRITE, compact ? SUBSCRIPTION_COMPACT : SUBSCRIPTION_NATURAL, true)};
    pmTaskDetails d(&c, sizeof(c), mem, 1, g_cb, 1);
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s == pmSuccess) {
        s = pmWaitForTaskCompletion(th);
        pmReleaseTask(th);
    }
    if (s == pmSuccess) {
        std::lock_guard<std::mutex> g(g_mtx);
        info[0] = g_rep.length;
        info[1] = g_rep.read_off;
        info[2] = g_rep.write_off;
        info[3] = (uint64_t)g_rep.has_read;
        info[4] = (uint64_t)g_rep.has_write;
        memcpy(view, g_rep.bytes.data(), g_rep.bytes.size() < view_cap ? g_rep.bytes.size() : view_cap);
    }
    pmFetchMemory(h);
    pmGetRawMemPtr(h, &p);
    memcpy(space, p, L);
    pmReleaseMemory(h);
    return (int)s;
}
