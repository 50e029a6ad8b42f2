// Public pm:: API of the runtime. Every entry point catches runtime errors and returns a
// pmStatus, as the reference does (pmPublicDefinitions.cpp:50-66); the message of the last
// failure is kept for pmGetLastError.
#include "pm_internal.h"
#include "unicorn_hip.h"

#include <climits>
#include <cstring>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace pm {

using namespace rt;

// ---------------------------------------------------------------- struct constructors
// Defaults follow the reference (pmPublicDefinitions.cpp:376-640).

pmSubscriptionInfo::pmSubscriptionInfo() : offset(0), length(0) {}
pmSubscriptionInfo::pmSubscriptionInfo(size_t o, size_t l) : offset(o), length(l) {}
pmScatteredSubscriptionInfo::pmScatteredSubscriptionInfo() : offset(0), size(0), step(0), count(0) {}
pmScatteredSubscriptionInfo::pmScatteredSubscriptionInfo(size_t o, size_t s, size_t st, size_t c)
    : offset(o), size(s), step(st), count(c) {}
pmGpuContext::pmGpuContext() : scratchBuffer(NULL), reservedGlobalMem(NULL) {}
pmSplitInfo::pmSplitInfo() : splitId(0), splitCount(0) {}
pmSplitInfo::pmSplitInfo(uint id, uint count) : splitId(id), splitCount(count) {}
pmMemInfo::pmMemInfo() : ptr(NULL), readPtr(NULL), writePtr(NULL), length(0), visibilityType(SUBSCRIPTION_NATURAL) {}
pmMemInfo::pmMemInfo(pmRawMemPtr p, pmRawMemPtr r, pmRawMemPtr w, size_t l)
    : ptr(p), readPtr(r), writePtr(w), length(l), visibilityType(SUBSCRIPTION_NATURAL) {}
pmMemInfo::pmMemInfo(pmRawMemPtr p, pmRawMemPtr r, pmRawMemPtr w, size_t l, pmSubscriptionVisibilityType v)
    : ptr(p), readPtr(r), writePtr(w), length(l), visibilityType(v) {}
pmSubtaskInfo::pmSubtaskInfo() : subtaskId(0), memCount(0) {}
pmSubtaskInfo::pmSubtaskInfo(ulong id, pmMemInfo* m, uint count) : subtaskId(id), memCount(count) {
    for (uint i = 0; i < count && i < MAX_MEM_SECTIONS_PER_TASK; ++i) memInfo[i] = m[i];
}
pmTaskInfo::pmTaskInfo() : taskHandle(NULL), taskConf(NULL), taskConfLength(0), taskId(0), subtaskCount(0), priority(0), originatingHost(0) {}
pmDataTransferInfo::pmDataTransferInfo() : memHandle(NULL), memLength(0), operatedMemLength(NULL), memType(MAX_MEM_TYPE), srcHost(0), destHost(0) {}
pmDeviceInfo::pmDeviceInfo() : deviceHandle(NULL), deviceType(MAX_DEVICE_TYPES), host(0), deviceIdOnHost(0), deviceIdInCluster(0) {
    name[0] = 0;
    description[0] = 0;
}
pmRedistributionMetadata::pmRedistributionMetadata() : order(0), count(0) {}
pmRedistributionMetadata::pmRedistributionMetadata(uint o, uint c) : order(o), count(c) {}

pmCallbacks::pmCallbacks()
    : dataDistribution(NULL), subtask_cpu(NULL), subtask_gpu_hip(NULL), subtask_gpu_custom(NULL),
      dataReduction(NULL), dataRedistribution(NULL), deviceSelection(NULL), preDataTransfer(NULL),
      postDataTransfer(NULL), taskCompletionCallback(NULL), subtask_opencl(NULL) {}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, const char* ocl) : pmCallbacks() {
    dataDistribution = d; subtask_opencl = ocl;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, const char* ocl, pmDataReductionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_opencl = ocl; dataReduction = r;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, const char* ocl, pmDataRedistributionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_opencl = ocl; dataRedistribution = r;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_HIP g) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_hip = g;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_Custom g) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_custom = g;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_HIP g, pmDataReductionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_hip = g; dataReduction = r;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_Custom g, pmDataReductionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_custom = g; dataReduction = r;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_HIP g, pmDataRedistributionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_hip = g; dataRedistribution = r;
}
pmCallbacks::pmCallbacks(pmDataDistributionCallback d, pmSubtaskCallback_CPU c, pmSubtaskCallback_GPU_Custom g, pmDataRedistributionCallback r) : pmCallbacks() {
    dataDistribution = d; subtask_cpu = c; subtask_gpu_custom = g; dataRedistribution = r;
}

pmGpuLaunchConf::pmGpuLaunchConf() : blocksX(1), blocksY(1), blocksZ(1), threadsX(1), threadsY(1), threadsZ(1), sharedMem(0) {}
pmGpuLaunchConf::pmGpuLaunchConf(int bx, int by, int bz, int tx, int ty, int tz)
    : blocksX(bx), blocksY(by), blocksZ(bz), threadsX(tx), threadsY(ty), threadsZ(tz), sharedMem(0) {}
pmGpuLaunchConf::pmGpuLaunchConf(int bx, int by, int bz, int tx, int ty, int tz, int sm)
    : blocksX(bx), blocksY(by), blocksZ(bz), threadsX(tx), threadsY(ty), threadsZ(tz), sharedMem(sm) {}

pmTaskMem::pmTaskMem() : memHandle(NULL), memType(MAX_MEM_TYPE), subscriptionVisibilityType(SUBSCRIPTION_NATURAL), disjointReadWritesAcrossSubtasks(false) {}
pmTaskMem::pmTaskMem(pmMemHandle h, pmMemType t) : memHandle(h), memType(t), subscriptionVisibilityType(SUBSCRIPTION_NATURAL), disjointReadWritesAcrossSubtasks(false) {}
pmTaskMem::pmTaskMem(pmMemHandle h, pmMemType t, pmSubscriptionVisibilityType v) : memHandle(h), memType(t), subscriptionVisibilityType(v), disjointReadWritesAcrossSubtasks(false) {}
pmTaskMem::pmTaskMem(pmMemHandle h, pmMemType t, pmSubscriptionVisibilityType v, bool d) : memHandle(h), memType(t), subscriptionVisibilityType(v), disjointReadWritesAcrossSubtasks(d) {}

pmTaskDetails::pmTaskDetails()
    : taskConf(NULL), taskConfLength(0), taskMem(NULL), taskMemCount(0), callbackHandle(NULL), subtaskCount(0),
      taskId(0), priority(0), policy(RANDOM_STEAL), timeOutInSecs(INT_MAX), multiAssignEnabled(true),
      overlapComputeCommunication(true), canSplitCpuSubtasks(false), canSplitGpuSubtasks(false),
      gpuCacheEnabled(true), suppressTaskLogs(false), affinityCriterion(MAXIMIZE_LOCAL_DATA), cluster(NULL) {}
pmTaskDetails::pmTaskDetails(void* c, uint cl, pmTaskMem* m, uint mc, pmCallbackHandle cb, ulong n) : pmTaskDetails() {
    taskConf = c; taskConfLength = cl; taskMem = m; taskMemCount = mc; callbackHandle = cb; subtaskCount = n;
}

// ---------------------------------------------------------------- error plumbing

namespace rt {
static std::mutex g_err_mtx;
static std::string g_last_error;
void set_last_error(const std::string& s) {
    std::lock_guard<std::mutex> g(g_err_mtx);
    g_last_error = s;
}
}  // namespace rt

#define PM_API_BEGIN try {
#define PM_API_END                                                   \
    }                                                                \
    catch (rt::Error & e) {                                          \
        rt::set_last_error(e.what);                                  \
        return e.status;                                             \
    }                                                                \
    catch (std::bad_alloc&) {                                        \
        rt::set_last_error("host allocation failed");                \
        return pmMemoryError;                                        \
    }                                                                \
    return pmSuccess;

static Runtime* need_runtime() {
    Runtime* r = Runtime::get();
    if (!r) raise(pmInitializationFailure, "pmInitialize has not been called");
    return r;
}

const char* pmGetLibVersion() { return "unicorn-mi355x_1_0_0"; }

const char* pmGetLastError() {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> g(rt::g_err_mtx);
    copy = rt::g_last_error;
    return copy.c_str();
}

pmStatus pmInitialize() {
    PM_API_BEGIN
    Runtime::init();
    PM_API_END
}

pmStatus pmFinalize() {
    PM_API_BEGIN
    Runtime::finalize();
    PM_API_END
}

uint pmGetHostId() { return 0; }
uint pmGetHostCount() { return 1; }

pmStatus pmRegisterCallbacks(const char* key, pmCallbacks cb, pmCallbackHandle* handle) {
    PM_API_BEGIN
    Runtime* r = need_runtime();
    if (!key || !handle) raise(pmInvalidKey, "null key or handle");
    if (strlen(key) >= MAX_CB_KEY_LEN) raise(pmMaxKeyLengthExceeded, "callback key too long");
    // reference pmPublicDefinitions.cpp:518-527
    if (cb.subtask_gpu_hip && cb.subtask_gpu_custom) raise(pmInvalidCallbacks, "both GPU callback kinds set");
    if (cb.subtask_opencl && (cb.subtask_cpu || cb.subtask_gpu_hip || cb.subtask_gpu_custom))
        raise(pmInvalidCallbacks, "OpenCL subtasks cannot be mixed with other kinds");
    if (cb.subtask_opencl) raise(pmInvalidCallbacks, "OpenCL subtasks are not supported");
    std::lock_guard<std::mutex> g(r->cb_mtx);
    if (r->callbacks.count(key)) raise(pmInvalidKey, "callback key already registered");
    std::unique_ptr<CallbackSet> set(new CallbackSet());
    set->key = key;
    set->cb = cb;
    *handle = set.get();
    r->callbacks[key] = std::move(set);
    PM_API_END
}

pmStatus pmReleaseCallbacks(pmCallbackHandle handle) {
    PM_API_BEGIN
    Runtime* r = need_runtime();
    std::lock_guard<std::mutex> g(r->cb_mtx);
    for (auto it = r->callbacks.begin(); it != r->callbacks.end(); ++it) {
        if (it->second.get() == handle) {
            r->callbacks.erase(it);
            return pmSuccess;
        }
    }
    raise(pmInvalidKey, "unknown callback handle");
    PM_API_END
}

pmStatus pmCreateMemory(size_t len, pmMemHandle* h) {
    PM_API_BEGIN
    need_runtime();
    if (!h) raise(pmFatalError, "null handle");
    *h = Memory::get().create(len, 0, 0);
    PM_API_END
}

pmStatus pmCreateMemory2D(size_t rows, size_t cols, pmMemHandle* h) {
    PM_API_BEGIN
    need_runtime();
    if (!h) raise(pmFatalError, "null handle");
    *h = Memory::get().create(rows * cols, rows, cols);
    PM_API_END
}

pmStatus pmReleaseMemory(pmMemHandle h) {
    PM_API_BEGIN
    need_runtime();
    if (!h) raise(pmUnrecognizedMemory, "null memory handle");
    Memory::get().release((AddressSpace*)h);
    PM_API_END
}

pmStatus pmFetchMemory(pmMemHandle h) {
    PM_API_BEGIN
    need_runtime();
    if (!h) raise(pmUnrecognizedMemory, "null memory handle");
    Memory::get().sync_host((AddressSpace*)h);
    PM_API_END
}

pmStatus pmFetchMemoryRange(pmMemHandle h, size_t off, size_t len) {
    PM_API_BEGIN
    need_runtime();
    AddressSpace* as = (AddressSpace*)h;
    if (!as) raise(pmUnrecognizedMemory, "null memory handle");
    if (off + len > as->length) raise(pmInvalidOffset, "range beyond address space");
    Memory::get().sync_host(as);
    PM_API_END
}

pmStatus pmGetRawMemPtr(pmMemHandle h, pmRawMemPtr* p) {
    PM_API_BEGIN
    need_runtime();
    AddressSpace* as = (AddressSpace*)h;
    if (!as || !p) raise(pmUnrecognizedMemory, "null memory handle");
    // The application may read and write through the pointer: make the host current and
    // treat any HBM mirror as stale from here on.
    Memory::get().sync_host(as);
    Memory::get().mark_host_written(as);
    *p = as->host;
    PM_API_END
}

static Subtask& ctx_subtask(pmTaskHandle th, ulong sid) {
    DistributionContext* c = current_context();
    if (!c || !c->subtask || !c->distribution) raise(pmInvalidCommand, "only callable from a dataDistribution callback");
    if ((Task*)th != c->task || sid != c->subtask->id) raise(pmInvalidIndex, "subscription for another subtask");
    return *c->subtask;
}

pmStatus pmSubscribeToMemory(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, uint mi,
                             pmSubscriptionType type, const pmSubscriptionInfo& info) {
    PM_API_BEGIN
    Subtask& st = ctx_subtask(th, sid);
    Task* t = (Task*)th;
    if (mi >= t->spaces.size()) raise(pmInvalidIndex, "memory index out of range");
    if (info.offset + info.length > t->spaces[mi]->length) raise(pmInvalidOffset, "subscription beyond address space");
    if (t->read_only(mi) && type != READ_SUBSCRIPTION) raise(pmInvalidCommand, "write subscription to read-only memory");
    if (t->write_only(mi) && type != WRITE_SUBSCRIPTION) raise(pmInvalidCommand, "read subscription to write-only memory");
    st.mem[mi].add(type, info.offset, info.length);
    PM_API_END
}

pmStatus pmSubscribeToMemory(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, uint mi,
                             pmSubscriptionType type, const pmScatteredSubscriptionInfo& info) {
    PM_API_BEGIN
    Subtask& st = ctx_subtask(th, sid);
    Task* t = (Task*)th;
    if (mi >= t->spaces.size()) raise(pmInvalidIndex, "memory index out of range");
    AddressSpace* as = t->spaces[mi];
    if (info.count && info.size) {
        if (info.step < info.size && info.count > 1) raise(pmInvalidOffset, "scattered step smaller than size");
        if (info.offset + (info.count - 1) * info.step + info.size > as->length)
            raise(pmInvalidOffset, "scattered subscription beyond address space");
        // reference pmSubscriptionManager.cpp:489-490: a scattered row may not exceed a 2-D row
        if (as->rows && info.size > as->cols) raise(pmFatalError, "scattered size exceeds the 2-D row length");
    }
    if (t->read_only(mi) && type != READ_SUBSCRIPTION) raise(pmInvalidCommand, "write subscription to read-only memory");
    if (t->write_only(mi) && type != WRITE_SUBSCRIPTION) raise(pmInvalidCommand, "read subscription to write-only memory");
    st.mem[mi].add_scattered(type, info);
    PM_API_END
}

// Registers a range of the subtask's output for redistribution (reference
// pmSubscriptionManager::RegisterRedistribution, pmSubscriptionManager.cpp:861-877): pOffset is
// relative to the subtask's consolidated write subscription of space pMemIndex. Only from the
// task's dataRedistribution callback.
pmStatus pmRedistributeData(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, uint mi, size_t off, size_t len,
                            uint order) {
    PM_API_BEGIN
    DistributionContext* c = current_context();
    if (!c || !c->redistribution || (Task*)th != c->task || c->subtask->id != sid)
        raise(pmInvalidCommand, "pmRedistributeData outside the subtask's dataRedistribution callback");
    Task* t = c->task;
    if (mi >= t->spaces.size()) raise(pmInvalidIndex, "memory index out of range");
    if (!len) return pmSuccess;
    const MemSubscription& ms = c->subtask->mem[mi];
    if (!ms.has_write) raise(pmInvalidCommand, "pmRedistributeData: the subtask has no write subscription on this space");
    const size_t global = ms.cwrite.offset + off;
    if (global >= t->spaces[mi]->length || len > t->spaces[mi]->length - global)
        raise(pmInvalidOffset, "pmRedistributeData: range beyond the address space");
    std::lock_guard<std::mutex> g(t->redist_mtx);
    auto& v = t->redist[mi];
    uint seq = 0;
    for (auto it = v.rbegin(); it != v.rend() && it->subtask == sid; ++it) ++seq;
    v.push_back({order, sid, seq, global, len});
    PM_API_END
}

pmStatus pmSetGpuLaunchConf(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, pmGpuLaunchConf& lc) {
    PM_API_BEGIN
    Subtask& st = ctx_subtask(th, sid);
    st.launch = lc;
    st.has_launch = true;
    PM_API_END
}

pmStatus pmReserveGpuGlobalMem(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, size_t size) {
    PM_API_BEGIN
    Subtask& st = ctx_subtask(th, sid);
    st.reserved_gpu = size;
    PM_API_END
}

pmStatus pmSubmitTask(pmTaskDetails d, pmTaskHandle* th) { return pmSubmitTaskWithOptions(d, nullptr, th); }

pmStatus pmSubmitTaskWithOptions(pmTaskDetails d, const pmTaskSubmitOptions* opt, pmTaskHandle* th) {
    PM_API_BEGIN
    Runtime* r = need_runtime();
    if (opt && opt->hipMode != -1 && (opt->hipMode < PM_HIP_MODE_PLAIN || opt->hipMode > PM_HIP_MODE_BATCHED))
        raise(pmInvalidCommand, "invalid HIP stub mode in the submit options");
    if (!th) raise(pmFatalError, "null task handle pointer");
    if (d.taskMemCount > MAX_MEM_SECTIONS_PER_TASK) raise(pmBeyondComputationalLimits, "too many memory sections");
    if (d.taskMemCount && !d.taskMem) raise(pmFatalError, "taskMem is null");
    CallbackSet* cs = (CallbackSet*)d.callbackHandle;
    {
        std::lock_guard<std::mutex> g(r->cb_mtx);
        bool found = false;
        for (auto& kv : r->callbacks) if (kv.second.get() == cs) found = true;
        if (!found) raise(pmInvalidKey, "unknown callback handle");
    }
    std::unique_ptr<Task> t(new Task());
    t->details = d;
    t->callbacks = cs;
    t->subtask_count = d.subtaskCount;
    if (d.taskConf && d.taskConfLength) {
        t->conf.assign((uint8_t*)d.taskConf, (uint8_t*)d.taskConf + d.taskConfLength);
    }
    for (uint i = 0; i < d.taskMemCount; ++i) {
        if (!d.taskMem[i].memHandle) raise(pmUnrecognizedMemory, "null memory handle in taskMem");
        if (d.taskMem[i].memType >= MAX_MEM_TYPE) raise(pmInvalidCommand, "invalid memory type");
        t->mem.push_back(d.taskMem[i]);
        t->spaces.push_back((AddressSpace*)d.taskMem[i].memHandle);
    }
    t->cpu_allowed = cs->cb.subtask_cpu != nullptr;
    t->gpu_allowed = r->gpu_count > 0 && (cs->cb.subtask_gpu_hip || cs->cb.subtask_gpu_custom);
    if (!t->cpu_allowed && !t->gpu_allowed) raise(pmNoCompatibleDevice, "no device can run this task's callbacks");
    t->hip_mode = opt && opt->hipMode != -1 ? opt->hipMode : r->hip_mode.load();
    t->unordered = opt && opt->unordered;
    // Splitting, as the reference's pmSubtaskSplitter: CPU splitting takes precedence over GPU
    // splitting (pmSubtaskSplitter.cpp:47-143). Subtasks of a task with a reduction are not
    // split here (the pieces would need one shared reduction buffer).
    if (!cs->cb.dataReduction) {
        if (d.canSplitCpuSubtasks && t->cpu_allowed) {
            const int f = r->split_factor_cpu.load();
            t->split_device = 1;
            t->split_count = (uint)(f > 0 ? f : (int)r->cpu_devices.size());
        } else if (d.canSplitGpuSubtasks && t->gpu_allowed) {
            t->split_device = 2;
            const int f = r->split_factor_gpu.load();
            t->split_count = (uint)(f > 0 ? f : r->gpu_count);
        }
        if (t->split_count < 2) t->split_device = 0;
    }
    // multi-assign (the reference's default, pmTaskDetails::multiAssignEnabled) needs both stub
    // kinds and whole subtasks; a reduced task's copies would need a second shadow per subtask
    t->multi_assign = d.multiAssignEnabled && t->cpu_allowed && t->gpu_allowed && !t->split_device &&
                      !cs->cb.dataReduction;
    for (uint i = 0; i < d.taskMemCount; ++i) t->has_deferred |= t->deferred(i);
    // held-back writes work on the host copy: such a task stages its GPU subtasks the PLAIN way
    // (a redistributing task runs resident and is redistributed in HBM)
    t->gpu_resident = t->gpu_allowed && !t->cpu_allowed && t->hip_mode != PM_HIP_MODE_PLAIN && !t->has_deferred;
    if (d.timeOutInSecs >= 0 && d.timeOutInSecs != INT_MAX) t->deadline = now_secs() + d.timeOutInSecs;
    for (uint i = 0; i < t->spaces.size(); ++i) {
        AddressSpace* as = t->spaces[i];
        std::lock_guard<std::mutex> g(as->mtx);
        as->users++;
    }
    // host copies are made current when the task becomes ready (Runtime::make_ready), after the
    // earlier tasks on its address spaces have completed
    Task* raw = t.release();
    *th = raw;
    if (raw->multi_assign) {
        std::lock_guard<std::mutex> g(r->ma_list_mtx);
        r->ma_tasks.push_back(raw);
    }
    r->submit(raw);
    PM_API_END
}

pmStatus pmWaitForTaskCompletion(pmTaskHandle th) {
    try {
        Runtime* r = need_runtime();
        Task* t = (Task*)th;
        if (!t) raise(pmFatalError, "null task handle");
        r->wait(t);
        return (pmStatus)t->status.load();
    } catch (rt::Error& e) {
        rt::set_last_error(e.what);
        return e.status;
    }
}

pmStatus pmReleaseTask(pmTaskHandle th) {
    pmStatus s = pmWaitForTaskCompletion(th);
    if (th) {
        Task* t = (Task*)th;
        {
            // multi-assigned copies that lost may still be running
            std::unique_lock<std::mutex> g(t->mtx);
            t->cv.wait(g, [&] { return t->runners == 0; });
        }
        if (t->dev_conf) rt::Memory::get().device_free(t->dev_conf, t->conf.size());
        delete t;
    }
    return s;
}

pmStatus pmGetTaskExecutionTimeInSecs(pmTaskHandle th, double* tm) {
    pmStatus s = pmWaitForTaskCompletion(th);
    if (th && tm) {
        Task* t = (Task*)th;
        *tm = t->end_time - t->submit_time;
    }
    return s;
}

pmStatus pmReleaseTaskAndResources(pmTaskDetails d, pmTaskHandle th) {
    pmStatus s = pmReleaseTask(th);
    if (d.callbackHandle) pmReleaseCallbacks(d.callbackHandle);
    for (uint i = 0; i < d.taskMemCount; ++i) pmReleaseMemory(d.taskMem[i].memHandle);
    return s;
}

void* pmGetScratchBufferHostFunc(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, pmScratchBufferType type, size_t size) {
    DistributionContext* c = current_context();
    if (!c || !c->subtask || (Task*)th != c->task || c->subtask->id != sid) return NULL;
    auto& buf = c->subtask->scratch[(int)type];
    if (buf.empty() && size) buf.assign(size, 0);
    return buf.empty() ? NULL : buf.data();
}

pmStatus pmReleaseScratchBuffer(pmTaskHandle th, pmDeviceHandle, ulong sid, pmSplitInfo&, pmScratchBufferType type) {
    DistributionContext* c = current_context();
    if (!c || !c->subtask || (Task*)th != c->task || c->subtask->id != sid) return pmInvalidCommand;
    c->subtask->scratch.erase((int)type);
    return pmSuccess;
}

void* pmGetLastReductionScratchBuffer(pmTaskHandle) { return NULL; }

// (order, bytes) of every order of a redistributed space, ascending (reference
// pmRedistributor::GetRedistributionMetadata, pmRedistributor.cpp:295-323); valid until the task
// is released.
pmRedistributionMetadata* pmGetRedistributionMetadata(pmTaskHandle th, uint mi, ulong* count) {
    if (count) *count = 0;
    Task* t = (Task*)th;
    if (!t || !t->finished.load(std::memory_order_acquire)) return NULL;
    std::lock_guard<std::mutex> g(t->redist_mtx);
    auto it = t->redist_meta.find(mi);
    if (it == t->redist_meta.end() || it->second.empty()) return NULL;
    if (count) *count = it->second.size();
    return it->second.data();
}

// ---------------------------------------------------------------- utilities: mapped files

namespace {
struct Mapping { void* addr; size_t len; };
std::mutex g_map_mtx;
std::map<std::string, Mapping> g_maps;
}

void* pmGetMappedFile(const char* path) {
    if (!path) return NULL;
    std::lock_guard<std::mutex> g(g_map_mtx);
    auto it = g_maps.find(path);
    return it == g_maps.end() ? NULL : it->second.addr;
}

pmStatus pmMapFile(const char* path) {
    if (!path || strlen(path) >= MAX_FILE_SIZE_LEN) return pmInvalidKey;
    std::lock_guard<std::mutex> g(g_map_mtx);
    if (g_maps.count(path)) return pmSuccess;
    int fd = open(path, O_RDONLY);
    if (fd < 0) { rt::set_last_error(std::string("cannot open ") + path); return pmConfFileNotFound; }
    struct stat sb;
    if (fstat(fd, &sb) != 0) { close(fd); return pmFatalError; }
    void* a = mmap(NULL, sb.st_size ? sb.st_size : 1, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (a == MAP_FAILED) return pmMemoryError;
    g_maps[path] = Mapping{a, (size_t)sb.st_size};
    return pmSuccess;
}

pmStatus pmUnmapFile(const char* path) {
    if (!path) return pmInvalidKey;
    std::lock_guard<std::mutex> g(g_map_mtx);
    auto it = g_maps.find(path);
    if (it == g_maps.end()) return pmInvalidKey;
    munmap(it->second.addr, it->second.len ? it->second.len : 1);
    g_maps.erase(it);
    return pmSuccess;
}

pmStatus pmMapFiles(const char* const* paths, uint n) {
    for (uint i = 0; i < n; ++i) {
        pmStatus s = pmMapFile(paths[i]);
        if (s != pmSuccess) return s;
    }
    return pmSuccess;
}

pmStatus pmUnmapFiles(const char* const* paths, uint n) {
    for (uint i = 0; i < n; ++i) {
        pmStatus s = pmUnmapFile(paths[i]);
        if (s != pmSuccess) return s;
    }
    return pmSuccess;
}

// ---------------------------------------------------------------- extensions

pmStatus pmSetHipStubMode(int mode) {
    if (mode < PM_HIP_MODE_PLAIN || mode > PM_HIP_MODE_BATCHED) return pmInvalidCommand;
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    r->hip_mode = mode;
    return pmSuccess;
}
int pmGetHipStubMode() { Runtime* r = Runtime::get(); return r ? r->hip_mode.load() : -1; }
pmStatus pmSetHipOptimalView(pmSubscriptionVisibilityType v) {
    if (v != SUBSCRIPTION_NATURAL && v != SUBSCRIPTION_COMPACT) return pmInvalidCommand;
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    r->hip_view = v;
    return pmSuccess;
}
pmStatus pmSetHipGraphs(bool e) {
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    r->use_graphs = e ? 1 : 0;
    return pmSuccess;
}
pmStatus pmSetHipPipelineDepth(int d) {
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    if (d < 1 || d > 16) return pmInvalidCommand;
    r->pipeline_depth = d;
    return pmSuccess;
}
pmStatus pmSetHipBatchLimit(int n) {
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    if (n < 1) return pmInvalidCommand;
    r->batch_limit = n;
    return pmSuccess;
}
pmStatus pmSetHbmBudget(size_t b) {
    Memory::get().set_budget(b ? b : Memory::get().default_budget);
    return pmSuccess;
}
size_t pmGetHbmBudget() { return Memory::get().budget(); }
size_t pmGetHbmResident() { return Memory::get().resident(); }

static int g_cpu_stub_request = 0;
pmStatus pmSetCpuStubCount(int n) {
    if (n < 0) return pmInvalidCommand;
    g_cpu_stub_request = n;
    setenv("PMLIB_MAX_CPU_PER_HOST", std::to_string(n).c_str(), 1);
    return pmSuccess;
}
pmStatus pmGetHipStubStats(pmHipStubStats* s) {
    Runtime* r = Runtime::get();
    if (!r || !s) return pmInitializationFailure;
    std::lock_guard<std::mutex> g(r->stats_mtx);
    *s = r->stats;
    return pmSuccess;
}
pmStatus pmSetSentinelCompressionThreshold(size_t bytes, size_t* previous) {
    const size_t prev = g_sentinel_threshold.exchange(bytes);
    if (previous) *previous = prev;
    return pmSuccess;
}
pmStatus pmResetHipStubStats() {
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    std::lock_guard<std::mutex> g(r->stats_mtx);
    r->stats = pmHipStubStats{};
    return pmSuccess;
}
pmStatus pmReadMemoryRange(pmMemHandle h, size_t off, size_t len, void* dst) {
    PM_API_BEGIN
    need_runtime();
    AddressSpace* as = (AddressSpace*)h;
    if (!as) raise(pmUnrecognizedMemory, "null memory handle");
    if (off > as->length || len > as->length - off) raise(pmInvalidOffset, "range beyond address space");
    if (len && !dst) raise(pmInvalidKey, "null destination");
    if (len) Memory::get().read_range(as, off, len, dst);
    PM_API_END
}

pmStatus pmPrefetchToDevice(pmMemHandle h) {
    PM_API_BEGIN
    Runtime* r = need_runtime();
    if (!r->gpu_count) raise(pmNoCompatibleDevice, "no HIP device");
    Memory::get().ensure_device((AddressSpace*)h, r->copy_stream);
    if (uh_stream_sync(r->copy_stream)) raise(pmGraphicsCardError, uh_last_error());
    PM_API_END
}
pmStatus pmGetDeviceMemPtr(pmMemHandle h, void** p) {
    PM_API_BEGIN
    Runtime* r = need_runtime();
    if (!r->gpu_count) raise(pmNoCompatibleDevice, "no HIP device");
    Memory::get().ensure_device((AddressSpace*)h, r->copy_stream);
    if (uh_stream_sync(r->copy_stream)) raise(pmGraphicsCardError, uh_last_error());
    *p = ((AddressSpace*)h)->dev;
    PM_API_END
}
pmStatus pmMarkDeviceWritten(pmMemHandle h) {
    PM_API_BEGIN
    need_runtime();
    Memory::get().mark_device_written((AddressSpace*)h);
    PM_API_END
}
bool pmHasHipDevice() { Runtime* r = Runtime::get(); return r && r->gpu_count > 0; }

pmStatus pmSetHipKernelTimingEvery(unsigned int every) {
    Runtime* r = Runtime::get();
    if (!r || !r->gpu_count) return pmNoCompatibleDevice;
    if (every > (unsigned)INT_MAX) return pmInvalidIndex;
    return uh_batch_set_timing(r->batch, (int)every) ? pmGraphicsCardError : pmSuccess;
}

pmStatus pmSetHipKernelTiming(bool on) { return pmSetHipKernelTimingEvery(on ? 1u : 0u); }

pmStatus pmGetHipKernelTime(double* ms, uint64_t* flushes, bool reset) {
    Runtime* r = Runtime::get();
    if (!r || !r->gpu_count) return pmNoCompatibleDevice;
    return uh_batch_timing(r->batch, ms, flushes, reset ? 1 : 0) ? pmGraphicsCardError : pmSuccess;
}

pmStatus pmSetSplitFactors(int pCpuPieces, int pGpuPieces) {
    Runtime* r = Runtime::get();
    if (!r) return pmInitializationFailure;
    if (pCpuPieces < 0 || pGpuPieces < 0 || pCpuPieces > 4096 || pGpuPieces > 4096) return pmInvalidCommand;
    r->split_factor_cpu = pCpuPieces;
    r->split_factor_gpu = pGpuPieces;
    return pmSuccess;
}

pmStatus pmGetHipKernelTimeByFamily(int family, double* ms, uint64_t* flushes, double* work, bool reset) {
    Runtime* r = Runtime::get();
    if (!r || !r->gpu_count) return pmNoCompatibleDevice;
    int rc = uh_batch_timing_family(r->batch, family, ms, flushes, work, reset ? 1 : 0);
    return rc == UH_ERR_ARG ? pmInvalidIndex : rc ? pmGraphicsCardError : pmSuccess;
}

}  // namespace pm
