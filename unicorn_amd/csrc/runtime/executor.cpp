// Tasks, per-device queues and the execution stubs of the single-node runtime.
//
// Reference structure (source/code/src): pmScheduler pushes subtask ranges to a priority queue
// per stub (pmExecutionStub.cpp:142-242, pmSafePriorityQueue.cpp), CPU stubs run subtask_cpu
// (pmExecutionStub.cpp:2638-2780), and the CUDA stub runs the per-subtask staging pipeline of
// pmExecutionStub.cpp:2167-2470 / 3216-4070. Here:
//   - every stub pulls subtask ids from the task at the head of one shared priority order
//     (priority, then submission order), which is the reference's PULL model without peers to
//     steal from on one node;
//   - CPU stubs: one thread per host core, one subtask at a time;
//   - the HIP stub: one thread for device 0 with three modes (pmExtensions.h): PLAIN follows the
//     reference stub (pinned gather, H2D into a COMPACT buffer, callback, D2H, scatter; a cache
//     of read-only buffers keyed by exact subscription), RESIDENT runs callbacks on views of HBM
//     mirrors, BATCHED additionally replaces per-subtask launches by a device work queue drained
//     by persistent kernels, and gathers subtasks of all active tasks into each batch.
#include "pm_internal.h"
#include "unicorn_hip.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <unordered_map>

namespace pm {
namespace rt {

Runtime* Runtime::inst_ = nullptr;

std::atomic<size_t> g_sentinel_threshold{(size_t)1 << 30};   // CUDA_SENTINEL_COMPRESSION_THRESHOLD

double now_secs() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Bounded spin before a thread sleeps on a condition variable: a client that submits its next
// task as soon as the previous one completes, and a waiter on a short task, are served without
// a futex wake-up (tens of microseconds). PMLIB_SPIN_US sets the bound (0 disables spinning).
static double spin_secs() {
    static const double s = [] {
        const char* e = getenv("PMLIB_SPIN_US");
        return (e ? atof(e) : 300.0) * 1e-6;
    }();
    return s;
}

// Subtasks of the first BATCHED flush onto an idle device (later flushes grow from it).
static constexpr unsigned long kFirstBatch = 128;

static inline void cpu_relax() {
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#endif
}

[[noreturn]] void raise(pmStatus s, const std::string& what) { throw Error{s, what}; }

namespace {
void stat_add(uint64_t pmHipStubStats::*f, uint64_t v);
}

pmTaskInfo Task::info() {
    pmTaskInfo ti;
    ti.taskHandle = (pmTaskHandle)this;
    ti.taskConf = conf.empty() ? nullptr : conf.data();
    ti.taskConfLength = (uint)conf.size();
    ti.taskId = details.taskId;
    ti.subtaskCount = subtask_count;
    ti.priority = details.priority;
    ti.originatingHost = 0;
    return ti;
}

bool Task::writable(uint i) const {
    pmMemType m = mem[i].memType;
    return m == WRITE_ONLY || m == READ_WRITE || m == WRITE_ONLY_LAZY || m == READ_WRITE_LAZY;
}
bool Task::read_only(uint i) const { return !writable(i); }
bool Task::write_only(uint i) const {
    pmMemType m = mem[i].memType;
    return m == WRITE_ONLY || m == WRITE_ONLY_LAZY;
}
bool Task::is_reduced(uint i) const {
    return callbacks->cb.dataReduction != nullptr && writable(i);
}
bool Task::deferred(uint i) const {
    return writable(i) && !write_only(i) && !is_reduced(i) && !mem[i].disjointReadWritesAcrossSubtasks;
}
void Task::defer_write(ulong subtask, uint m, size_t offset, const void* src, size_t len) {
    DeferredWrite w;
    w.subtask = subtask;
    w.mem = m;
    w.offset = offset;
    w.bytes.assign((const uint8_t*)src, (const uint8_t*)src + len);
    std::lock_guard<std::mutex> g(deferred_mtx);
    deferred_writes.push_back(std::move(w));
}
pmSubscriptionVisibilityType Task::visibility(uint i, bool gpu) const {
    pmSubscriptionVisibilityType v = mem[i].subscriptionVisibilityType;
    if (v == SUBSCRIPTION_OPTIMAL) {
        if (!gpu) return SUBSCRIPTION_NATURAL;
        Runtime* r = Runtime::get();
        if (!gpu_resident) return SUBSCRIPTION_COMPACT;   // staged like the reference stub
        return (pmSubscriptionVisibilityType)r->hip_view.load();
    }
    return v;
}

bool Task::has_work(bool gpu) {
    if (gpu ? !gpu_allowed : !cpu_allowed) return false;
    if (next.load() < subtask_count) return true;
    if (split_device != (gpu ? 2 : 1)) return false;
    std::lock_guard<std::mutex> g(split_mtx);
    return !open_splits.empty();
}

bool Task::claim_commit(ulong sid, bool gpu_copy) {
    std::lock_guard<std::mutex> g(ma_mtx);
    auto it = cpu_runs.find(sid);
    if (it == cpu_runs.end()) return !gpu_copy;   // a HIP copy after the CPU run ended: it lost
    if (it->second.committed) return false;
    it->second.committed = true;
    return true;
}

Runtime* Runtime::get() { return inst_; }

static int env_int(const char* name, int dflt) {
    const char* v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

void Runtime::init() {
    if (inst_) return;
    Runtime* r = new Runtime();
    inst_ = r;
    int n = 0;
    if (!getenv("PMLIB_DISABLE_HIP") && uh_device_count(&n) == 0 && n > 0) {
        r->gpu_count = 1;
        uh_set_device(0);
        pmDeviceInfo* d = new pmDeviceInfo();
        d->deviceHandle = d;
        size_t total = 0;
        int cus = 0;
        uh_device_props(0, d->name, MAX_NAME_STR_LEN, &total, &cus, nullptr, nullptr, nullptr);
        snprintf(d->description, MAX_DESC_STR_LEN, "HIP device 0, %d CUs, %zu MiB HBM", cus, total >> 20);
        d->deviceType = GPU_HIP;
        d->deviceIdOnHost = 0;
        d->deviceIdInCluster = 0;
        r->gpu_device = d;
        size_t freeb = 0, totb = 0;
        uh_mem_info(&freeb, &totb);
        size_t budget = (size_t)((double)freeb * 0.85);
        int mb = env_int("PMLIB_HBM_BUDGET_MB", 0);
        if (mb > 0) budget = (size_t)mb << 20;
        Memory::get().set_budget(budget);
        Memory::get().default_budget = budget;
        for (int i = 0; i < 16; ++i) {
            uh_stream_create(&r->gpu_streams[i]);
            uh_event_create(&r->gpu_events[i]);
        }
        uh_stream_create(&r->batch_stream);
        uh_stream_create(&r->copy_stream);
        uh_batch_create(&r->batch, r->batch_stream);
        // device copies handed to kernel-kind callbacks: the stub's pmDeviceInfo and one status
        // word per subtask slot (16) and for the batch
        if (uh_malloc(&r->dev_info, sizeof(pmDeviceInfo)) ||
            uh_memcpy_h2d(r->dev_info, d, sizeof(pmDeviceInfo), r->copy_stream) ||
            uh_malloc(&r->dev_status, (kBatchStatusSlot + 1) * sizeof(pmStatus)) ||
            uh_stream_sync(r->copy_stream))
            fprintf(stderr, "pmlib: device copies for kernel-kind callbacks: %s\n", uh_last_error());
    }
    r->hip_mode = env_int("PMLIB_HIP_MODE", r->hip_mode.load());
    r->use_graphs = env_int("PMLIB_HIP_GRAPHS", 1);
    int ncpu = r->cpu_workers.load();
    if (!ncpu) ncpu = env_int("PMLIB_MAX_CPU_PER_HOST", (int)std::thread::hardware_concurrency());
    if (ncpu <= 0) ncpu = 1;
    for (int i = 0; i < ncpu; ++i) {
        pmDeviceInfo* d = new pmDeviceInfo();
        d->deviceHandle = d;
        snprintf(d->name, MAX_NAME_STR_LEN, "CPU core %d", i);
        snprintf(d->description, MAX_DESC_STR_LEN, "host CPU stub %d", i);
        d->deviceType = CPU;
        d->deviceIdOnHost = (uint)(i + r->gpu_count);
        d->deviceIdInCluster = d->deviceIdOnHost;
        r->cpu_devices.push_back(d);
    }
    for (int i = 0; i < ncpu; ++i) r->cpu_threads.emplace_back(&Runtime::cpu_loop, r, i);
    if (r->gpu_count) r->gpu_thread = std::thread(&Runtime::gpu_loop, r);
}

void Runtime::finalize() {
    Runtime* r = inst_;
    if (!r) return;
    {
        std::lock_guard<std::mutex> g(r->q_mtx);
        r->stopping = true;
    }
    r->q_cv.notify_all();
    for (auto& t : r->cpu_threads) t.join();
    if (r->gpu_thread.joinable()) r->gpu_thread.join();
    if (r->gpu_count) {
        uh_device_sync();
        uh_batch_destroy(r->batch);
        for (int i = 0; i < 16; ++i) {
            uh_stream_destroy(r->gpu_streams[i]);
            uh_event_destroy(r->gpu_events[i]);
        }
        uh_stream_destroy(r->batch_stream);
        uh_stream_destroy(r->copy_stream);
        uh_free(r->dev_info);
        uh_free(r->dev_status);
    }
    for (auto* d : r->cpu_devices) delete d;
    delete r->gpu_device;
    inst_ = nullptr;
    delete r;
}

// Set while this thread holds q_mtx inside pick / cancel_locked: a task completed there releases
// its dependents later (drain_ready), since making them ready takes q_mtx.
static thread_local bool tl_queue_locked = false;

struct QueueLockedScope {
    bool saved;
    QueueLockedScope() : saved(tl_queue_locked) { tl_queue_locked = true; }
    ~QueueLockedScope() { tl_queue_locked = saved; }
};

// Ordering. Like the reference, which lets one task at a time lock an address space and queues
// the others (pmAddressSpace::EnqueueForLock / ScanLockQueue, pmAddressSpace.cpp:500-530;
// pmTask::LockAddressSpaces, pmTask.cpp:176-193), a task does not start before the earlier
// tasks it conflicts with have completed. Conflict is per address space: either task writes it
// (tasks that only read a space, such as served requests on one dataset, run together). A task
// submitted as unordered (pmSubmitTaskWithOptions) skips the wait; later tasks still wait for it.
void Runtime::submit(Task* t) {
    t->submit_time = now_secs();
    {
        std::lock_guard<std::mutex> g(q_mtx);
        t->seq = ++seq;
        if (t->subtask_count == 0) {
            t->state = TaskState::Done;
            t->finished.store(true, std::memory_order_release);
            t->start_time = t->end_time = t->submit_time;
            return;
        }
    }
    bool ready_now;
    {
        std::lock_guard<std::mutex> g(dep_mtx);
        std::vector<std::pair<AddressSpace*, bool>> uses;     // distinct spaces, written or not
        for (uint i = 0; i < t->spaces.size(); ++i) {
            auto it = std::find_if(uses.begin(), uses.end(), [&](auto& u) { return u.first == t->spaces[i]; });
            if (it == uses.end()) uses.push_back({t->spaces[i], t->writable(i)});
            else it->second = it->second || t->writable(i);
        }
        std::vector<Task*> deps;
        for (auto& u : uses) {
            if (!t->unordered)
                for (auto& h : u.first->holders)
                    if ((h.second || u.second) && std::find(deps.begin(), deps.end(), h.first) == deps.end())
                        deps.push_back(h.first);
            u.first->holders.push_back({t, u.second});
        }
        for (Task* d : deps) d->dependents.push_back(t);
        t->pending = (int)deps.size();
        ready_now = t->pending == 0;
    }
    if (ready_now) make_ready(t);
}

void Runtime::make_ready(Task* t) {
    // Subtasks that read or write host memory need the host copy current. This runs once the
    // earlier conflicting tasks have completed, so it sees their results.
    if (t->cpu_allowed || !t->gpu_resident) {
        for (uint i = 0; i < t->spaces.size(); ++i) {
            try {
                Memory::get().sync_host(t->spaces[i]);
                if (t->writable(i)) Memory::get().mark_host_written(t->spaces[i]);
            } catch (Error& e) {
                int expected = pmSuccess;
                t->status.compare_exchange_strong(expected, (int)e.status);
                set_last_error(e.what);
            }
        }
    }
    {
        std::lock_guard<std::mutex> g(q_mtx);
        // keep `active` ordered by (priority, seq): lower priority value first
        auto it = std::find_if(active.begin(), active.end(), [&](Task* o) {
            return o->details.priority > t->details.priority ||
                   (o->details.priority == t->details.priority && o->seq > t->seq);
        });
        active.insert(it, t);
    }
    submitted.fetch_add(1, std::memory_order_release);
    q_cv.notify_all();
}

void Runtime::drain_ready() {
    std::vector<Task*> ready;
    {
        std::lock_guard<std::mutex> g(dep_mtx);
        ready.swap(deferred_ready);
    }
    for (Task* t : ready) make_ready(t);
}

void Runtime::wait(Task* t) {
    if (!t->finished.load(std::memory_order_acquire)) {
        const double until = now_secs() + spin_secs();
        while (!t->finished.load(std::memory_order_acquire) && now_secs() < until) cpu_relax();
    }
    std::unique_lock<std::mutex> g(t->mtx);
    if (t->state == TaskState::Done) return;
    t->cv.wait(g, [&] { return t->state == TaskState::Done; });
    stat_add(&pmHipStubStats::ns_wake, (uint64_t)((now_secs() - t->end_time) * 1e9));
}

Task* Runtime::pick(bool gpu, ulong& first, ulong& count, ulong max_count, pmSplitInfo* split) {
    // caller holds q_mtx
    QueueLockedScope locked;
    if (split) *split = pmSplitInfo();
    double now = 0;
    for (auto it = active.begin(); it != active.end(); ++it) {
        Task* t = *it;
        if (gpu ? !t->gpu_allowed : !t->cpu_allowed) continue;
        if (gpu && t->gpu_resident && !t->gpu_prepared) continue;   // not admitted yet
        if (t->deadline > 0) {
            if (now == 0) now = now_secs();
            if (now > t->deadline) {
                // time-out: remaining subtasks are cancelled (reference TASK_TIME_OUT,
                // pmTimedEventManager.cpp:92-95)
                ulong f = t->next.exchange(t->subtask_count);
                if (f < t->subtask_count || t->split_device) {
                    int expected = pmSuccess;
                    t->status.compare_exchange_strong(expected, (int)pmTimerFailure);
                    active.erase(it);
                    drop_open_splits(t);
                    if (f < t->subtask_count) finish_subtasks(t, t->subtask_count - f);
                    return pick(gpu, first, count, max_count, split);
                }
            }
        }
        if (t->split_device == (gpu ? 2 : 1) && split) {
            // one piece: the next split of an open subtask, or split 0 of a new one
            std::lock_guard<std::mutex> sg(t->split_mtx);
            if (!t->open_splits.empty()) {
                auto& rec = t->open_splits.front();
                first = rec.first;
                *split = pmSplitInfo(rec.second, t->split_count);
                if (++rec.second == t->split_count) t->open_splits.pop_front();
            } else {
                const ulong f = t->next.load();
                if (f >= t->subtask_count) continue;
                t->next = f + 1;
                first = f;
                *split = pmSplitInfo(0, t->split_count);
                t->open_splits.push_back({f, 1});
                t->split_left[f] = t->split_count;
            }
            count = 1;
            if (t->next.load() >= t->subtask_count && t->open_splits.empty()) active.erase(it);
        } else {
            ulong f = t->next.load();
            if (f >= t->subtask_count) continue;
            ulong n = std::min(max_count, t->subtask_count - f);
            t->next = f + n;
            if (t->next.load() >= t->subtask_count) {
                std::lock_guard<std::mutex> sg(t->split_mtx);
                if (t->open_splits.empty()) active.erase(it);
            }
            first = f;
            count = n;
        }
        if (t->state == TaskState::Queued) {
            t->state = TaskState::Running;
            t->start_time = now_secs();
            stat_add(&pmHipStubStats::ns_queue, (uint64_t)((t->start_time - t->submit_time) * 1e9));
        }
        return t;
    }
    return nullptr;
}

void Runtime::cancel_locked(Task* t, pmStatus status) {
    QueueLockedScope locked;
    ulong f = t->next.exchange(t->subtask_count);
    bool open;
    {
        std::lock_guard<std::mutex> sg(t->split_mtx);
        open = !t->open_splits.empty();
    }
    if (f >= t->subtask_count && !open) return;
    int expected = pmSuccess;
    t->status.compare_exchange_strong(expected, (int)status);
    auto it = std::find(active.begin(), active.end(), t);
    if (it != active.end()) active.erase(it);
    drop_open_splits(t);
    if (f < t->subtask_count) finish_subtasks(t, t->subtask_count - f);
}

void Runtime::drop_open_splits(Task* t) {
    ulong done = 0;
    {
        std::lock_guard<std::mutex> sg(t->split_mtx);
        for (auto& rec : t->open_splits) {
            uint& left = t->split_left[rec.first];
            left -= t->split_count - rec.second;      // pieces never handed out
            if (left == 0) {
                t->split_left.erase(rec.first);
                ++done;
            }
        }
        t->open_splits.clear();
    }
    if (done) finish_subtasks(t, done);
}

void Runtime::finish_piece(Task* t, ulong subtask, const pmSplitInfo& split) {
    if (split.splitCount) {
        std::lock_guard<std::mutex> sg(t->split_mtx);
        auto it = t->split_left.find(subtask);
        if (it == t->split_left.end() || --it->second > 0) return;
        t->split_left.erase(it);
    }
    finish_subtasks(t, 1);
}

void Runtime::runner_enter(Task* t) {
    {
        std::lock_guard<std::mutex> g(t->mtx);
        t->runners++;
    }
    for (AddressSpace* as : t->spaces) {
        std::lock_guard<std::mutex> g(as->mtx);
        as->runners++;
    }
}

void Runtime::runner_exit(Task* t) {
    for (AddressSpace* as : t->spaces) {
        std::lock_guard<std::mutex> g(as->mtx);
        if (--as->runners == 0) as->runners_cv.notify_all();
    }
    // last touch of the task: once runners is 0 its owner may delete it
    std::lock_guard<std::mutex> g(t->mtx);
    if (--t->runners == 0) t->cv.notify_all();
}

// Picks a subtask that a CPU stub is executing for the idle HIP stub to copy (reference
// pmExecutionStub.cpp:1082-1143). Rates are this task's: the HIP stub's time per subtask Tg
// (wall time from its first staging to its last retirement over the subtasks it finished) and
// a CPU stub's mean subtask time Tc. A copy starts when 1.05 * Tg is below the CPU's expected
// remaining time Tc - elapsed (SUBTASK_TRANSFER_OVERHEAD); before any CPU subtask has finished,
// once the CPU run has taken 0.95 * Tg (MA_WAIT_FACTOR_LENIENT, a GPU asking a CPU on the same
// host). A HIP stub that has finished no subtask of the task has no rate and copies nothing, as
// in the reference. retry_in: seconds after which a waiting run may qualify (0: none).
Task* Runtime::ma_candidate(ulong& sid, double& retry_in) {
    retry_in = 0;
    std::lock_guard<std::mutex> lg(ma_list_mtx);
    const double now = now_secs();
    for (Task* t : ma_tasks) {
        if (t->status.load() != pmSuccess) continue;
        std::lock_guard<std::mutex> g(t->ma_mtx);
        if (!t->gpu_finished || t->cpu_runs.empty()) continue;
        const double tg = (t->gpu_last - t->gpu_first) / (double)t->gpu_finished;
        const double tc = t->cpu_finished ? t->cpu_secs / (double)t->cpu_finished : 0.0;
        for (auto& kv : t->cpu_runs) {
            Task::CpuRun& run = kv.second;
            if (run.duplicated || run.committed) continue;
            const double elapsed = now - run.start;
            bool go;
            if (tc == 0.0) {
                go = elapsed >= 0.95 * tg;
                if (!go) {
                    const double w = 0.95 * tg - elapsed;
                    retry_in = retry_in == 0 ? w : std::min(retry_in, w);
                }
            } else {
                go = 1.05 * tg < tc - elapsed;
            }
            if (!go) continue;
            run.duplicated = true;
            sid = kv.first;
            {
                std::lock_guard<std::mutex> tm(t->mtx);
                t->runners++;
            }
            for (AddressSpace* as : t->spaces) {
                std::lock_guard<std::mutex> am(as->mtx);
                as->runners++;
            }
            return t;
        }
    }
    return nullptr;
}

void Runtime::finish_subtasks(Task* t, ulong n) {
    ulong d = t->done.fetch_add(n) + n;
    if (d == t->subtask_count) complete(t);
}

// ---------------------------------------------------------------- reductions

// Completed subtasks of a reduced task are folded pairwise in completion order into one
// accumulated shadow (reference pmReducer::AddSubtask, pmReducer.cpp:191-217).
static void reduce_into(Task* t, std::unique_ptr<Subtask> st) {
    std::lock_guard<std::mutex> g(t->reduce_mtx);
    if (!t->pending_reduce) {
        t->pending_reduce = std::move(st);
        return;
    }
    Subtask* acc = t->pending_reduce.get();
    pmReductionOpType op;
    pmReductionDataType dt;
    bool builtin = is_builtin_reduction(t->callbacks->cb.dataReduction, &op, &dt);
    for (uint i = 0; i < t->spaces.size(); ++i) {
        if (!t->is_reduced(i) || !st->shadow[i]) continue;
        if (acc->shadow_on_device[i] != st->shadow_on_device[i])
            raise(pmFatalError, "mixed host/device reduction shadows");
        if (acc->shadow_on_device[i]) {
            if (!builtin) raise(pmFatalError, "HBM shadows take built-in reductions only");
            static const size_t kElem[] = {4, 4, 8, 8, 4, 8};    // pmReductionDataType sizes
            if (uh_reduce(acc->shadow[i], st->shadow[i], acc->shadow_len[i] / kElem[dt], (int)op, (int)dt,
                          Runtime::get()->copy_stream) ||
                uh_stream_sync(Runtime::get()->copy_stream))
                raise(pmGraphicsCardError, uh_last_error());
            Memory::get().device_free(st->shadow[i], st->shadow_len[i]);
            st->shadow[i] = nullptr;
        } else if (builtin) {
            reduce_host(acc->shadow[i], st->shadow[i], acc->shadow_len[i], op, dt);
        }
    }
    if (!builtin) {
        // user-defined reduction: both subtasks are described by NATURAL views of their shadows
        auto view = [&](Subtask* s) {
            pmSubtaskInfo si;
            si.subtaskId = s->id;
            si.memCount = (uint)t->spaces.size();
            for (uint i = 0; i < t->spaces.size(); ++i) {
                if (t->is_reduced(i) && s->shadow[i]) {
                    si.memInfo[i].ptr = si.memInfo[i].writePtr = s->shadow[i];
                    si.memInfo[i].length = s->shadow_len[i];
                }
            }
            return si;
        };
        pmDeviceInfo d1 = *(acc->device ? acc->device : Runtime::get()->cpu_devices[0]);
        pmDeviceInfo d2 = *(st->device ? st->device : Runtime::get()->cpu_devices[0]);
        pmStatus s = t->callbacks->cb.dataReduction(t->info(), d1, view(acc), d2, view(st.get()));
        if (s != pmSuccess) raise(s, "dataReduction callback failed");
    }
    for (uint i = 0; i < t->spaces.size(); ++i)
        if (st->shadow[i] && !st->shadow_on_device[i]) { free(st->shadow[i]); st->shadow[i] = nullptr; }
}

// Write the accumulated shadow of every reduced space back into the address space.
static void commit_reduction(Task* t) {
    std::lock_guard<std::mutex> g(t->reduce_mtx);
    Subtask* acc = t->pending_reduce.get();
    if (!acc) return;
    for (uint i = 0; i < t->spaces.size(); ++i) {
        if (!t->is_reduced(i) || !acc->shadow[i]) continue;
        AddressSpace* as = t->spaces[i];
        pmSubscriptionInfo u = acc->mem[i].unified();
        if (acc->shadow_on_device[i]) {
            Memory::get().ensure_device(as, Runtime::get()->copy_stream);
            if (uh_memcpy_d2d(as->dev + u.offset, acc->shadow[i], std::min(acc->shadow_len[i], span_in_space(as, u)),
                              Runtime::get()->copy_stream) ||
                uh_stream_sync(Runtime::get()->copy_stream))
                raise(pmGraphicsCardError, uh_last_error());
            Memory::get().mark_device_written(as);
            Memory::get().device_free(acc->shadow[i], acc->shadow_len[i]);
        } else {
            Memory::get().sync_host(as);
            memcpy(as->host + u.offset, acc->shadow[i], std::min(acc->shadow_len[i], span_in_space(as, u)));
            Memory::get().mark_host_written(as);
            free(acc->shadow[i]);
        }
        acc->shadow[i] = nullptr;
    }
    t->pending_reduce.reset();
}

void Runtime::complete(Task* t) {
    if (t->multi_assign) {
        std::lock_guard<std::mutex> g(ma_list_mtx);
        ma_tasks.erase(std::remove(ma_tasks.begin(), ma_tasks.end(), t), ma_tasks.end());
    }
    if (t->has_deferred) {
        // held-back writes land in subtask order (stable within a subtask)
        std::lock_guard<std::mutex> g(t->deferred_mtx);
        auto& w = t->deferred_writes;
        std::stable_sort(w.begin(), w.end(), [](const Task::DeferredWrite& a, const Task::DeferredWrite& b) {
            return a.subtask < b.subtask;
        });
        if (t->status.load() == pmSuccess)
            for (auto& x : w) memcpy(t->spaces[x.mem]->host + x.offset, x.bytes.data(), x.bytes.size());
        std::vector<Task::DeferredWrite>().swap(w);
    }
    try {
        commit_reduction(t);
        if (t->status.load() == pmSuccess && t->callbacks->cb.dataRedistribution) commit_redistribution(t);
    } catch (Error& e) {
        int expected = pmSuccess;
        t->status.compare_exchange_strong(expected, (int)e.status);
        set_last_error(e.what);
    }
    const bool pinned = t->gpu_resident && t->gpu_prepared;
    for (AddressSpace* as : t->spaces) {
        std::lock_guard<std::mutex> g(as->mtx);
        as->users--;
        if (pinned) as->gpu_pins--;
    }
    if (pinned) gpu_admitted--;
    std::vector<Task*> ready;
    {
        std::lock_guard<std::mutex> g(dep_mtx);
        for (AddressSpace* as : t->spaces) {
            auto& h = as->holders;
            h.erase(std::remove_if(h.begin(), h.end(), [&](auto& x) { return x.first == t; }), h.end());
        }
        for (Task* d : t->dependents)
            if (--d->pending == 0) {
                if (tl_queue_locked) deferred_ready.push_back(d);
                else ready.push_back(d);
            }
        t->dependents.clear();
    }
    for (Task* d : ready) make_ready(d);
    bool copies_running;
    {
        std::lock_guard<std::mutex> g(t->mtx);
        copies_running = t->runners > 0;
    }
    // every kernel that read the device conf has finished once its subtasks retired; a
    // multi-assigned copy still running keeps it (pmReleaseTask frees it after the copy ends)
    if (t->dev_conf && !copies_running) {
        Memory::get().device_free(t->dev_conf, t->conf.size());
        t->dev_conf = nullptr;
    }
    if (t->callbacks->cb.taskCompletionCallback) t->callbacks->cb.taskCompletionCallback(t->info());
    std::lock_guard<std::mutex> g(t->mtx);
    t->end_time = now_secs();
    t->state = TaskState::Done;
    t->finished.store(true, std::memory_order_release);
    t->cv.notify_all();
}

static void fail_task(Task* t, const Error& e) {
    int expected = pmSuccess;
    t->status.compare_exchange_strong(expected, (int)e.status);
    set_last_error(e.what);
}

// ---------------------------------------------------------------- CPU stubs

void Runtime::cpu_loop(int idx) {
    pmDeviceInfo* dev = cpu_devices[idx];
    for (;;) {
        Task* t;
        ulong first = 0, count = 0;
        pmSplitInfo split;
        {
            std::unique_lock<std::mutex> g(q_mtx);
            q_cv.wait(g, [&] {
                if (stopping) return true;
                for (Task* x : active)
                    if (x->has_work(false)) return true;
                return false;
            });
            if (stopping) return;
            t = pick(false, first, count, 1, &split);
        }
        drain_ready();      // tasks released by a time-out inside pick
        if (!t) continue;
        for (ulong s = first; s < first + count; ++s) {
            std::unique_ptr<Subtask> st(new Subtask());
            st->id = s;
            st->split = split;
            const bool ma = t->multi_assign;
            double t0 = 0;
            if (ma) {
                runner_enter(t);
                t0 = now_secs();
                std::lock_guard<std::mutex> g(t->ma_mtx);
                t->cpu_runs[s].start = t0;
            }
            bool commit = true, claimed = false;
            try {
                if (t->status.load() == pmSuccess) {
                    if (t->callbacks->cb.deviceSelection &&
                        !t->callbacks->cb.deviceSelection(t->info(), *dev))
                        raise(pmNoCompatibleDevice, "device selection rejected the CPU stub");
                    run_distribution(t, *st, dev);
                    std::vector<std::vector<uint8_t>> compact;
                    pmSubtaskInfo si = host_subtask_info(t, *st, compact, ma);
                    pmStatus rs;
                    {
                        SubtaskScope scope(t, st.get(), dev->deviceHandle);
                        rs = t->callbacks->cb.subtask_cpu(t->info(), *dev, si);
                    }
                    if (rs != pmSuccess) raise(rs, "subtask_cpu callback failed");
                    // a multi-assigned subtask writes the address space only if this copy is first
                    if (ma) {
                        commit = t->claim_commit(s, false);
                        claimed = true;
                    }
                    if (commit) {
                        for (uint i = 0; i < t->spaces.size(); ++i) {
                            if (!t->writable(i) || t->is_reduced(i)) continue;
                            const bool hold = t->deferred(i);
                            auto put = [&](size_t off, const uint8_t* src, size_t len) {
                                if (hold) t->defer_write(s, i, off, src, len);
                                else memcpy(t->spaces[i]->host + off, src, len);
                            };
                            if (si.memInfo[i].visibilityType == SUBSCRIPTION_COMPACT) {
                                // COMPACT views: scatter written records back
                                CompactLayout cl = make_compact(st->mem[i]);
                                for (auto& kv : st->mem[i].write().map())
                                    put(kv.first, compact[i].data() + cl.compact_offset(kv.first), kv.second);
                            } else if (st->shadow[i]) {
                                // private NATURAL view: copy its written records
                                const size_t base = st->mem[i].unified().offset;
                                for (auto& kv : st->mem[i].write().map())
                                    put(kv.first, (uint8_t*)st->shadow[i] + (kv.first - base), kv.second);
                            }
                        }
                    }
                    if (commit && t->callbacks->cb.dataRedistribution) run_redistribution(t, *st, dev);
                    if (t->callbacks->cb.dataReduction) reduce_into(t, std::move(st));
                }
            } catch (Error& e) {
                if (ma && !claimed) {
                    commit = t->claim_commit(s, false);
                    claimed = true;
                }
                if (commit) fail_task(t, e);
            }
            if (ma && !claimed) commit = t->claim_commit(s, false);
            if (st && !t->callbacks->cb.dataReduction)
                for (uint i = 0; i < MAX_MEM_SECTIONS_PER_TASK; ++i)
                    if (st->shadow[i] && !st->shadow_on_device[i]) { free(st->shadow[i]); st->shadow[i] = nullptr; }
            if (ma) {
                std::lock_guard<std::mutex> g(t->ma_mtx);
                t->cpu_runs.erase(s);
                if (commit) {
                    t->cpu_secs += now_secs() - t0;
                    t->cpu_finished++;
                }
            }
            if (commit) finish_piece(t, s, split);
            if (ma) runner_exit(t);
        }
    }
}

// ---------------------------------------------------------------- HIP stub

namespace {

void hipchk(int rc, const char* what) {
    if (rc) raise(pmGraphicsCardError, std::string(what) + ": " + uh_last_error());
}

// One GPU subtask in flight.
struct GpuSlot {
    Task* task = nullptr;
    uint64_t seq = 0;              // staging order (the OOM back-off retires the oldest first)
    std::unique_ptr<Subtask> st;
    void* stream = nullptr;
    void* event = nullptr;
    // device buffers owned by this subtask (freed at completion)
    std::vector<std::pair<void*, size_t>> dev_bufs;
    // PLAIN: written records to scatter from the pinned buffer at completion
    // sent >= 0: the pinned bytes hold the record sentinel-compressed, `sent` non-zero elements
    // of sent_elem bytes (uint64 count, the elements, their uint32 indices)
    struct Scatter { uint mem; size_t src_off, len, pinned_off; int64_t sent = -1; uint32_t sent_elem = 0; };
    std::vector<Scatter> scatters;
    uint8_t* pinned = nullptr;
    size_t pinned_cap = 0;
    // RESIDENT + COMPACT: written records to scatter device->mirror at completion
    struct DevScatter { AddressSpace* as; void* buf; CompactLayout cl; };
    std::vector<DevScatter> dev_scatters;
    // PLAIN: read-only cache entries this subtask reads (pinned until it completes)
    std::vector<std::pair<uint64_t, std::vector<size_t>>> cache_refs;
    bool plain = false;     // staged the PLAIN way (host gather / scatter)
    bool busy = false;
    bool ma_copy = false;   // a multi-assigned copy of a CPU stub's subtask
    bool copy_failed = false;
    void* scratch_dev = nullptr;   // device copy of the subtask's scratch buffer
    int scratch_type = -1;
};

// Read-only device buffers of the PLAIN mode, keyed by (address space, exact records)
// (reference pmCudaCacheKey, pmDataTypes.h:295-387; LRU eviction).
struct CacheEntry {
    void* buf = nullptr;
    size_t bytes = 0;
    uint64_t last_use = 0;
    int users = 0;
    void* ready = nullptr;   // event recorded after the upload; later users wait on it
};

struct GpuState {
    GpuSlot slots[16];
    std::map<std::pair<uint64_t, std::vector<size_t>>, CacheEntry> cache;
    uint64_t clock = 0;
};

GpuState& gstate() {
    static GpuState s;
    return s;
}

void cache_purge(AddressSpace* as, bool all) {
    auto& c = gstate().cache;
    for (auto it = c.begin(); it != c.end();) {
        if ((all || (as && it->first.first == as->uid)) && it->second.users == 0) {
            Memory::get().device_free(it->second.buf, it->second.bytes);
            if (it->second.ready) uh_event_destroy(it->second.ready);
            it = c.erase(it);
        } else {
            ++it;
        }
    }
}

void cache_purge_uid(uint64_t uid) {
    auto& c = gstate().cache;
    for (auto it = c.begin(); it != c.end();) {
        if (it->first.first == uid && it->second.users == 0) {
            Memory::get().device_free(it->second.buf, it->second.bytes);
            if (it->second.ready) uh_event_destroy(it->second.ready);
            it = c.erase(it);
        } else {
            ++it;
        }
    }
}

void cache_shrink(size_t need) {
    auto& c = gstate().cache;
    while (Memory::get().resident() + need > Memory::get().budget()) {
        auto victim = c.end();
        for (auto it = c.begin(); it != c.end(); ++it)
            if (it->second.users == 0 && (victim == c.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == c.end()) return;
        Memory::get().device_free(victim->second.buf, victim->second.bytes);
        if (victim->second.ready) uh_event_destroy(victim->second.ready);
        c.erase(victim);
    }
}

// Out-of-memory back-off of the per-subtask path (reference pmExecutionStub.cpp:2403-2415: on a
// failed device allocation the stub waits for its oldest in-flight subtask and retries): while
// staging a subtask, the stub thread installs a hook that retires its oldest other in-flight
// subtask, which frees that subtask's buffers; it returns false when none is left.
static thread_local std::function<bool()>* tl_oom_backoff = nullptr;

void* slot_device_alloc(GpuSlot& s, size_t bytes) {
    if (!bytes) return nullptr;
    for (;;) {
        cache_shrink(bytes);
        try {
            void* p = Memory::get().device_alloc(bytes);
            s.dev_bufs.push_back({p, bytes});
            return p;
        } catch (Error& e) {
            if (e.status != pmMemoryError || !tl_oom_backoff || !(*tl_oom_backoff)()) throw;
            stat_add(&pmHipStubStats::oom_waits, 1);
        }
    }
}

uint8_t* slot_pinned(GpuSlot& s, size_t bytes) {
    if (bytes > s.pinned_cap) {
        if (s.pinned) uh_host_free(s.pinned);
        s.pinned = nullptr;
        size_t cap = std::max(bytes, (size_t)1 << 20);
        hipchk(uh_host_alloc((void**)&s.pinned, cap), "pinned staging allocation");
        s.pinned_cap = cap;
    }
    return s.pinned;
}

void stat_add(uint64_t pmHipStubStats::*f, uint64_t v) {
    Runtime* r = Runtime::get();
    std::lock_guard<std::mutex> g(r->stats_mtx);
    r->stats.*f += v;
}

// Copies a small host table to the device on the copy stream and waits for that copy only (not
// for the work queued on the subtask's or the batch's stream). The source stays host memory the
// caller owns; the slot's pinned buffer is not used, since it may still feed an H2D copy in flight.
void upload_small(void* dst, const void* src, size_t bytes, const char* what) {
    Runtime* r = Runtime::get();
    hipchk(uh_memcpy_h2d(dst, src, bytes, r->copy_stream), what);
    hipchk(uh_stream_sync(r->copy_stream), what);
}

// The subtask's scratch buffer, allocated (and possibly filled) on the host by its
// data-distribution callback, becomes the device buffer gpuContext.scratchBuffer (what
// pmGetScratchBuffer returns in device code). A buffer that outlives the subtask
// (SUBTASK_TO_POST_SUBTASK, PRE_SUBTASK_TO_POST_SUBTASK) is copied back after the callback's
// work (download_scratch), for the dataRedistribution callback and other post-subtask readers.
void upload_scratch(Subtask& st, GpuSlot& s, pmSubtaskInfo& si) {
    s.scratch_dev = nullptr;
    s.scratch_type = -1;
    for (int type : {(int)PRE_SUBTASK_TO_SUBTASK, (int)PRE_SUBTASK_TO_POST_SUBTASK, (int)SUBTASK_TO_POST_SUBTASK}) {
        auto it = st.scratch.find(type);
        if (it == st.scratch.end() || it->second.empty()) continue;
        void* d = slot_device_alloc(s, it->second.size());
        upload_small(d, it->second.data(), it->second.size(), "scratch upload");
        si.gpuContext.scratchBuffer = d;
        s.scratch_dev = d;
        s.scratch_type = type;
        return;
    }
}

void download_scratch(Subtask& st, GpuSlot& s, void* stream) {
    if (!s.scratch_dev || s.scratch_type == (int)PRE_SUBTASK_TO_SUBTASK) return;
    std::vector<uint8_t>& h = st.scratch[s.scratch_type];
    hipchk(uh_memcpy_d2h(h.data(), s.scratch_dev, h.size(), stream), "scratch download");
}

// PLAIN mode: stage every address space of the subtask as the reference stub does.
pmSubtaskInfo stage_plain(Task* t, GpuSlot& s) {
    Subtask& st = *s.st;
    pmSubtaskInfo si;
    si.subtaskId = st.id;
    si.memCount = (uint)t->spaces.size();
    si.splitInfo = st.split;
    // gather everything to upload into one pinned buffer
    struct Up { void* dst; size_t pin_off, len; uint mem; };
    std::vector<std::pair<uint64_t, std::vector<size_t>>> fresh;   // cache entries uploaded here
    std::vector<Up> ups;
    std::vector<std::pair<size_t, std::pair<const uint8_t*, size_t>>> gathers;  // pin_off <- src,len
    size_t pin_total = 0;
    size_t scatter_total = 0;
    struct PendingScatter { uint mem; size_t src_off, len; void* dev; };
    std::vector<PendingScatter> pend;
    for (uint i = 0; i < t->spaces.size(); ++i) {
        AddressSpace* as = t->spaces[i];
        MemSubscription& ms = st.mem[i];
        pmMemInfo mi;
        mi.visibilityType = t->visibility(i, true);
        const bool natural = mi.visibilityType == SUBSCRIPTION_NATURAL;
        // records to stage and their device positions
        std::vector<size_t> rsrc, rdst, rlen;
        size_t total = 0, base_off = 0;
        CompactLayout cl;
        if (natural) {
            pmSubscriptionInfo u = t->read_only(i) ? ms.cread : ms.unified();
            base_off = u.offset;
            total = u.length;
            for (auto& kv : ms.read().map()) { rsrc.push_back(kv.first); rdst.push_back(kv.first - base_off); rlen.push_back(kv.second); }
        } else {
            cl = make_compact(ms);
            total = cl.total;
            for (auto& kv : ms.read().map()) { rsrc.push_back(kv.first); rdst.push_back(cl.compact_offset(kv.first)); rlen.push_back(kv.second); }
        }
        if (!total) { si.memInfo[i] = mi; continue; }
        uint8_t* dbuf = nullptr;
        if (t->read_only(i) && t->details.gpuCacheEnabled) {
            std::vector<size_t> key;
            key.push_back(natural ? 1 : 0);
            key.push_back((size_t)as->version.load());
            for (auto& kv : ms.read().map()) { key.push_back(kv.first); key.push_back(kv.second); }
            auto& cache = gstate().cache;
            auto ck = std::make_pair(as->uid, key);
            auto it = cache.find(ck);
            if (it != cache.end()) {
                it->second.last_use = ++gstate().clock;
                it->second.users++;
                dbuf = (uint8_t*)it->second.buf;
                stat_add(&pmHipStubStats::cache_hits, 1);
                rsrc.clear();
                // the upload may still be in flight on another subtask's stream
                hipchk(uh_stream_wait_event(s.stream, it->second.ready), "cache wait");
            } else {
                cache_shrink(total);
                CacheEntry e;
                e.buf = Memory::get().device_alloc(total);
                e.bytes = total;
                e.users = 1;
                e.last_use = ++gstate().clock;
                hipchk(uh_event_create(&e.ready), "cache event");
                cache[ck] = e;
                dbuf = (uint8_t*)e.buf;
                fresh.push_back(ck);
            }
            s.cache_refs.push_back(ck);
        } else {
            dbuf = (uint8_t*)slot_device_alloc(s, total);
        }
        for (size_t r = 0; r < rsrc.size(); ++r) {
            gathers.push_back({pin_total, {as->host + rsrc[r], rlen[r]}});
            ups.push_back({dbuf + rdst[r], pin_total, rlen[r], i});
            pin_total += rlen[r];
        }
        if (t->writable(i)) {
            for (auto& kv : ms.write().map()) {
                size_t doff = natural ? kv.first - base_off : cl.compact_offset(kv.first);
                pend.push_back({i, kv.first, kv.second, dbuf + doff});
                scatter_total += kv.second;
            }
        }
        mi.ptr = dbuf;
        mi.length = total;
        if (natural) {
            if (t->read_only(i)) mi.readPtr = dbuf;
            else if (t->write_only(i)) mi.writePtr = dbuf;
            else {
                if (ms.has_read) mi.readPtr = dbuf + (ms.cread.offset - base_off);
                if (ms.has_write) mi.writePtr = dbuf + (ms.cwrite.offset - base_off);
            }
        } else {
            if (cl.has_read) mi.readPtr = dbuf + cl.first_read;
            if (cl.has_write) mi.writePtr = dbuf + cl.first_write;
        }
        if (t->is_reduced(i)) {
            st.shadow[i] = nullptr;   // plain mode reduces on the host after D2H
        }
        si.memInfo[i] = mi;
    }
    // host gather into pinned memory, then H2D (pmExecutionStub.cpp:3655-3750, 3370-3615)
    uint8_t* pin = slot_pinned(s, std::max(pin_total, scatter_total));
    for (auto& g : gathers) memcpy(pin + g.first, g.second.first, g.second.second);
    // one copy per run of records that is contiguous both in pinned memory and on the device
    // (a COMPACT read-only view becomes a single H2D, as pmCudaInterface.cu:482-487)
    for (size_t u = 0; u < ups.size();) {
        size_t v = u + 1, len = ups[u].len;
        while (v < ups.size() && ups[v].mem == ups[u].mem && ups[v].pin_off == ups[u].pin_off + len &&
               (uint8_t*)ups[v].dst == (uint8_t*)ups[u].dst + len) {
            len += ups[v].len;
            ++v;
        }
        hipchk(uh_memcpy_h2d(ups[u].dst, pin + ups[u].pin_off, len, s.stream), "H2D staging");
        u = v;
    }
    for (auto& ck : fresh) hipchk(uh_event_record(gstate().cache[ck].ready, s.stream), "cache event");
    stat_add(&pmHipStubStats::h2d_bytes, pin_total);
    // the D2H commands are issued after the callback (see run_plain)
    s.scatters.clear();
    size_t poff = 0;
    for (auto& p : pend) {
        s.scatters.push_back({p.mem, p.src_off, p.len, poff});
        poff += p.len;
    }
    // stash device sources of the scatters in dev_scatters' place
    s.dev_scatters.clear();
    for (size_t k = 0; k < pend.size(); ++k) {
        GpuSlot::DevScatter d;
        d.as = t->spaces[pend[k].mem];
        d.buf = pend[k].dev;
        s.dev_scatters.push_back(d);
    }
    return si;
}

// RESIDENT / BATCHED: views of HBM mirrors. COMPACT views are gathered on the device.
pmSubtaskInfo stage_resident(Task* t, GpuSlot& s, void* stream) {
    Subtask& st = *s.st;
    pmSubtaskInfo si;
    si.subtaskId = st.id;
    si.memCount = (uint)t->spaces.size();
    si.splitInfo = st.split;
    s.dev_scatters.clear();
    for (uint i = 0; i < t->spaces.size(); ++i) {
        AddressSpace* as = t->spaces[i];
        MemSubscription& ms = st.mem[i];
        pmMemInfo mi;
        mi.visibilityType = t->visibility(i, true);
        if (mi.visibilityType == SUBSCRIPTION_NATURAL) {
            pmSubscriptionInfo u = t->read_only(i) ? ms.cread : ms.unified();
            if (u.length) {
                uint8_t* base = as->dev + u.offset;
                if (t->is_reduced(i)) {
                    void* sh = Memory::get().device_alloc(u.length);
                    hipchk(uh_memcpy_d2d(sh, as->dev + u.offset, span_in_space(as, u), stream), "shadow init");
                    st.shadow[i] = sh;
                    st.shadow_len[i] = u.length;
                    st.shadow_on_device[i] = true;
                    base = (uint8_t*)sh;
                }
                mi.ptr = base;
                mi.length = u.length;
                if (t->read_only(i)) mi.readPtr = base;
                else if (t->write_only(i)) mi.writePtr = base;
                else {
                    if (ms.has_read) mi.readPtr = base + (ms.cread.offset - u.offset);
                    if (ms.has_write) mi.writePtr = base + (ms.cwrite.offset - u.offset);
                }
            }
        } else {
            if (t->is_reduced(i)) raise(pmFatalError, "COMPACT views of reduced spaces are not supported on the HIP stub");
            CompactLayout cl = make_compact(ms);
            if (cl.total) {
                uint8_t* buf = (uint8_t*)slot_device_alloc(s, cl.total);
                // record tables: src offsets, dst offsets, lengths
                size_t n = cl.src_off.size();
                std::vector<uint64_t> tab(3 * n);
                for (size_t r = 0; r < n; ++r) { tab[r] = cl.src_off[r]; tab[n + r] = cl.dst_off[r]; tab[2 * n + r] = cl.len[r]; }
                uint64_t* dtab = (uint64_t*)slot_device_alloc(s, tab.size() * 8);
                upload_small(dtab, tab.data(), tab.size() * 8, "record table upload");
                uint64_t maxlen = 0;
                for (size_t r = 0; r < n; ++r) maxlen = std::max<uint64_t>(maxlen, cl.len[r]);
                if (t->read_only(i) || !t->write_only(i)) {
                    hipchk(uh_copy_records(buf, as->dev, dtab, dtab + n, dtab + 2 * n, (uint32_t)n, maxlen, stream), "device gather");
                    stat_add(&pmHipStubStats::d2d_bytes, cl.total);
                }
                mi.ptr = buf;
                mi.length = cl.total;
                if (cl.has_read) mi.readPtr = buf + cl.first_read;
                if (cl.has_write) mi.writePtr = buf + cl.first_write;
                if (t->writable(i)) {
                    GpuSlot::DevScatter d;
                    d.as = as;
                    d.buf = buf;
                    d.cl = cl;
                    s.dev_scatters.push_back(d);
                }
            }
        }
        si.memInfo[i] = mi;
    }
    return si;
}

// After a RESIDENT/BATCHED subtask with COMPACT writable views: scatter written records back
// into the mirror (device to device), all records of a view in one copy kernel (a copy call per
// record costs a few microseconds of launch each: 2048 per 2048-row block).
void scatter_resident(Task* t, GpuSlot& s, void* stream) {
    for (auto& d : s.dev_scatters) {
        MemSubscription* ms = nullptr;
        for (uint i = 0; i < t->spaces.size(); ++i) if (t->spaces[i] == d.as) ms = &s.st->mem[i];
        if (!ms) continue;
        const auto& recs = ms->write().map();
        const size_t n = recs.size();
        if (!n) continue;
        std::vector<uint64_t> tab(3 * n);
        uint64_t maxlen = 0, bytes = 0;
        size_t r = 0;
        for (auto& kv : recs) {
            tab[r] = d.cl.compact_offset(kv.first);   // source: packed position in the view
            tab[n + r] = kv.first;                    // destination: offset in the address space
            tab[2 * n + r] = kv.second;
            maxlen = std::max<uint64_t>(maxlen, kv.second);
            bytes += kv.second;
            ++r;
        }
        uint64_t* dtab = (uint64_t*)slot_device_alloc(s, tab.size() * 8);
        upload_small(dtab, tab.data(), tab.size() * 8, "record table upload");
        hipchk(uh_copy_records(d.as->dev, d.buf, dtab, dtab + n, dtab + 2 * n, (uint32_t)n, maxlen, stream),
               "device scatter");
        stat_add(&pmHipStubStats::d2d_bytes, bytes);
    }
}

}  // namespace

// Launch a kernel-kind callback (reference pmCudaInterface.cu:532-543). taskInfo.taskConf points
// at a device copy of the task's configuration, uploaded once per task at its first kernel-kind
// subtask (reference CreateTaskConf, pmExecutionStub.cpp:3321-3337) and freed when the task
// completes; pDeviceInfo is the stub's device copy made at initialisation (reference
// CreateDeviceInfoCudaPtr, 3344-3368). The kernel writes its pmStatus into a word of the slot,
// which is not read back (as the reference, whose read-back is disabled at 3487-3497). Nothing
// here waits for the device: the launch is ordered after the subtask's staging on its stream.
// A subtask that set no launch configuration runs the default one (one block of one thread), as
// the reference's default pmCudaLaunchConf.
static pmStatus launch_kernel_kind(Task* t, Subtask& st, pmSubtaskInfo& si, void* stream, int status_slot) {
    Runtime* r = Runtime::get();
    pmTaskInfo ti = t->info();
    if (!t->conf.empty()) {
        if (!t->dev_conf) {
            void* d = Memory::get().device_alloc(t->conf.size());
            if (uh_memcpy_h2d(d, t->conf.data(), t->conf.size(), r->copy_stream) || uh_stream_sync(r->copy_stream)) {
                Memory::get().device_free(d, t->conf.size());
                raise(pmGraphicsCardError, std::string("taskConf upload: ") + uh_last_error());
            }
            t->dev_conf = d;
        }
        ti.taskConf = t->dev_conf;
    }
    pmDeviceInfo* ddev = (pmDeviceInfo*)r->dev_info;
    pmStatus* dstatus = (pmStatus*)r->dev_status + status_slot;
    void* args[] = {&ti, &ddev, &si, &dstatus};
    const pmGpuLaunchConf& lc = st.launch;
    hipchk(uh_launch_kernel((const void*)t->callbacks->cb.subtask_gpu_hip, lc.blocksX, lc.blocksY, lc.blocksZ,
                            lc.threadsX, lc.threadsY, lc.threadsZ, (size_t)lc.sharedMem, stream, args),
           "kernel-kind launch");
    return pmSuccess;
}

static pmStatus run_gpu_callback(Task* t, Subtask& st, pmSubtaskInfo& si, void* stream,
                                 pmDeviceInfo* dev, int status_slot) {
    if (t->callbacks->cb.subtask_gpu_custom) {
        SubtaskScope scope(t, &st, dev->deviceHandle);
        return t->callbacks->cb.subtask_gpu_custom(t->info(), *dev, si, stream);
    }
    return launch_kernel_kind(t, st, si, stream, status_slot);
}

static void release_slot(GpuSlot& s) {
    for (auto& b : s.dev_bufs) Memory::get().device_free(b.first, b.second);
    s.dev_bufs.clear();
    for (auto& k : s.cache_refs) {
        auto it = gstate().cache.find(k);
        if (it != gstate().cache.end() && it->second.users > 0) it->second.users--;
    }
    s.cache_refs.clear();
    s.scatters.clear();
    s.dev_scatters.clear();
    s.busy = false;
    s.task = nullptr;
    s.st.reset();
    s.ma_copy = s.copy_failed = false;
}

// Completion of one PLAIN/RESIDENT slot (its event has fired).
static void retire_slot(Runtime* r, GpuSlot& s, bool plain) {
    Task* t = s.task;
    const ulong sid = s.st->id;
    const pmSplitInfo split = s.st->split;
    // a multi-assigned copy writes the address space only if it finished first
    const bool copy = s.ma_copy;
    const bool win = !copy || (!s.copy_failed && t->claim_commit(sid, true));
    try {
        if (plain && win) {
            // scatter written records from pinned memory into the host address space
            // (reference CopyDataFromPinnedBuffers, pmExecutionStub.cpp:3752-3847); records of
            // reduced spaces go to a host shadow of the subtask instead
            for (auto& sc : s.scatters) {
                if (t->is_reduced(sc.mem)) {
                    pmSubscriptionInfo u = s.st->mem[sc.mem].unified();
                    if (!s.st->shadow[sc.mem]) {
                        s.st->shadow[sc.mem] = calloc(1, u.length);
                        if (!s.st->shadow[sc.mem]) raise(pmMemoryError, "shadow allocation");
                        s.st->shadow_len[sc.mem] = u.length;
                        s.st->shadow_on_device[sc.mem] = false;
                    }
                    uint8_t* dst = (uint8_t*)s.st->shadow[sc.mem] + (sc.src_off - u.offset);
                    if (sc.sent >= 0) {
                        // expand the sentinel-compressed record (reference
                        // UncompressSentinelCompressedData, pmExecutionStub.cpp:3883-3950)
                        const uint8_t* p = s.pinned + sc.pinned_off;
                        const uint64_t cnt = (uint64_t)sc.sent;
                        const size_t es = sc.sent_elem;
                        const uint8_t* vals = p + 8;
                        const uint32_t* idx = (const uint32_t*)(vals + cnt * es);
                        memset(dst, 0, sc.len);
                        for (uint64_t j = 0; j < cnt; ++j) {
                            if ((size_t)idx[j] * es + es > sc.len) raise(pmFatalError, "sentinel index out of range");
                            memcpy(dst + (size_t)idx[j] * es, vals + j * es, es);
                        }
                    } else {
                        memcpy(dst, s.pinned + sc.pinned_off, sc.len);
                    }
                } else if (t->deferred(sc.mem)) {
                    t->defer_write(sid, sc.mem, sc.src_off, s.pinned + sc.pinned_off, sc.len);
                } else {
                    memcpy(t->spaces[sc.mem]->host + sc.src_off, s.pinned + sc.pinned_off, sc.len);
                }
            }
        }
        if (win && t->callbacks->cb.dataRedistribution && t->status.load() == pmSuccess)
            run_redistribution(t, *s.st, r->gpu_device);
        if (t->callbacks->cb.dataReduction && t->status.load() == pmSuccess) {
            reduce_into(t, std::move(s.st));
        }
    } catch (Error& e) {
        fail_task(t, e);
    }
    if (t->multi_assign && win) {
        std::lock_guard<std::mutex> g(t->ma_mtx);
        t->gpu_last = now_secs();
        t->gpu_finished++;
    }
    stat_add(&pmHipStubStats::subtasks, 1);
    if (copy) stat_add(&pmHipStubStats::ma_copies, 1);
    if (copy && win) stat_add(&pmHipStubStats::ma_wins, 1);
    release_slot(s);
    if (win) r->finish_piece(t, sid, split);
    if (copy) r->runner_exit(t);
}

void Runtime::gpu_loop() {
    uh_set_device(0);
    GpuState& gs = gstate();
    int inflight = 0;
    int next_slot = 0;
    // batches in flight: event + list of (task, subtask) slots
    struct BatchRun {
        void* event = nullptr;
        std::vector<std::unique_ptr<GpuSlot>> slots;
    };
    std::deque<BatchRun> batches;
    // Subtasks one BATCHED flush may take. Onto an idle device the first flush is small, so the
    // kernels start while the host stages the rest; each later flush may take twice the one in
    // flight (staging a subtask costs the host less than the device spends on it, so the device
    // stays busy), up to the batch limit.
    size_t last_batch = 0;
    auto batch_cap = [&]() -> ulong {
        const ulong lim = (ulong)batch_limit.load();
        const ulong grow = batches.empty() ? kFirstBatch : std::max<ulong>(kFirstBatch, 2 * (ulong)last_batch);
        return std::min(lim, grow);
    };
    void* batch_events[4];
    for (int i = 0; i < 4; ++i) uh_event_create(&batch_events[i]);
    int batch_ev = 0;
    auto ensure_task_ready = [&](Task* t) {
        if (t->gpu_prepared) return;
        if (t->gpu_resident) {
            bool copied = false;
            for (AddressSpace* as : t->spaces) copied |= Memory::get().ensure_device(as, copy_stream);
            if (copied) hipchk(uh_stream_sync(copy_stream), "mirror upload");
            for (AddressSpace* as : t->spaces) {
                std::lock_guard<std::mutex> g(as->mtx);
                as->gpu_pins++;
            }
            gpu_admitted++;
        }
        t->gpu_prepared = true;
    };

    // Admission of GPU-resident tasks, in queue order: a task is admitted (its mirrors made
    // current and pinned) once its address spaces fit the HBM budget, evicting mirrors that no
    // admitted task uses. A task that does not fit holds back the tasks behind it until admitted
    // work completes; with nothing admitted it can never fit and fails with pmMemoryError.
    auto admit_queued = [&]() {
        std::vector<Task*> cand;
        {
            std::lock_guard<std::mutex> g(q_mtx);
            for (Task* x : active)
                if (x->gpu_allowed && x->gpu_resident && !x->gpu_prepared && x->has_work(true))
                    cand.push_back(x);
        }
        for (Task* x : cand) {
            bool ok = false;
            try {
                ok = Memory::get().admit(x->spaces);
                if (!ok) {
                    // idle read-only buffers of the PLAIN mode count against the budget too
                    cache_shrink(Memory::get().missing_bytes(x->spaces));
                    ok = Memory::get().admit(x->spaces);
                }
                if (ok) ensure_task_ready(x);
            } catch (Error& e) {
                fail_task(x, e);
                std::lock_guard<std::mutex> g(q_mtx);
                cancel_locked(x, e.status);
                continue;
            }
            if (!ok) {
                if (gpu_admitted.load() == 0) {
                    set_last_error("task address spaces exceed the HBM budget");
                    std::lock_guard<std::mutex> g(q_mtx);
                    cancel_locked(x, pmMemoryError);
                    continue;
                }
                break;
            }
        }
    };

    // Retires the finished batches in flight, oldest first; with wait_oldest, waits for the oldest
    // one first (only that one: the next flush is staged while the younger batches run).
    auto retire_batches = [&](bool wait_oldest) {
        while (!batches.empty()) {
            BatchRun& b = batches.front();
            if (wait_oldest) {
                const double w0 = now_secs();
                uh_event_sync(b.event);
                stat_add(&pmHipStubStats::ns_wait, (uint64_t)((now_secs() - w0) * 1e9));
                wait_oldest = false;
            } else if (uh_event_query(b.event) != 0) return;
            const double r0 = now_secs();
            for (auto& sp : b.slots) {
                GpuSlot& s = *sp;
                Task* t = s.task;
                try {
                    // the batch has finished (its event fired), so the scatter goes to the copy
                    // stream: waiting on the batch stream would also wait for the younger batch
                    if (!s.dev_scatters.empty()) {
                        scatter_resident(t, s, copy_stream);
                        hipchk(uh_stream_sync(copy_stream), "device scatter");
                    }
                } catch (Error& e) { fail_task(t, e); }
                retire_slot(this, s, false);
            }
            batches.pop_front();
            stat_add(&pmHipStubStats::ns_retire, (uint64_t)((now_secs() - r0) * 1e9));
        }
    };

    for (;;) {
        // retire finished per-subtask slots
        for (int i = 0; i < 16; ++i) {
            GpuSlot& s = gs.slots[i];
            if (s.busy && uh_event_query(s.event) == 0) {
                const bool plain = s.plain;
                if (!plain && !s.dev_scatters.empty()) {
                    try { scatter_resident(s.task, s, s.stream); hipchk(uh_stream_sync(s.stream), "scatter"); }
                    catch (Error& e) { fail_task(s.task, e); }
                }
                retire_slot(this, s, plain);
                inflight--;
            }
        }
        retire_batches(false);
        for (uint64_t uid : Memory::get().take_released()) cache_purge_uid(uid);
        admit_queued();
        drain_ready();      // tasks released by a cancellation or time-out under the queue lock
        const int depth = std::max(1, std::min(16, pipeline_depth.load()));
        static thread_local uint64_t slot_seq = 0;
        Task* t = nullptr;
        ulong first = 0, count = 0;
        pmSplitInfo split;
        int mode = PM_HIP_MODE_PLAIN;
        bool ma_copy = false;
        {
            std::unique_lock<std::mutex> g(q_mtx);
            auto pending = [&]() -> bool {
                for (Task* x : active)
                    if (x->has_work(true)) return true;
                return false;
            };
            // first task whose subtasks can be staged now (admitted, or staged per subtask)
            auto front = [&]() -> Task* {
                for (Task* x : active)
                    if (x->has_work(true) && (!x->gpu_resident || x->gpu_prepared))
                        return x;
                return nullptr;
            };
            if (inflight == 0 && batches.empty()) {
                if (!stopping && !pending()) {
                    // idle: run a copy of a subtask a CPU stub is executing (multi-assign)
                    g.unlock();
                    double retry = 0;
                    ulong sid = 0;
                    Task* mt = ma_candidate(sid, retry);
                    g.lock();
                    if (mt) {
                        t = mt;
                        first = sid;
                        count = 1;
                        ma_copy = true;
                    } else if (retry > 0) {
                        if (!stopping && !pending())
                            q_cv.wait_for(g, std::chrono::microseconds((long)(retry * 1e6) + 1));
                        continue;
                    }
                }
                if (!ma_copy) {
                    if (!stopping && !pending() && spin_secs() > 0) {
                        const uint64_t seen = submitted.load(std::memory_order_acquire);
                        g.unlock();
                        const double until = now_secs() + spin_secs();
                        while (submitted.load(std::memory_order_acquire) == seen && now_secs() < until) cpu_relax();
                        g.lock();
                    }
                    q_cv.wait(g, [&] { return stopping || pending(); });
                    if (stopping && !pending()) break;
                    if (!front()) continue;   // admission runs first
                }
            } else if (!front() || inflight >= depth || (int)batches.size() >= 2) {
                g.unlock();
                // wait for the oldest work to finish
                if (!batches.empty()) retire_batches(true);
                else {
                    for (int i = 0; i < 16; ++i)
                        if (gs.slots[i].busy) { uh_event_sync(gs.slots[i].event); break; }
                }
                continue;
            }
            if (!ma_copy) {
                // the mode is the one the task was submitted under
                mode = front()->hip_mode;
                ulong max_take = (mode == PM_HIP_MODE_BATCHED && front()->gpu_resident) ? batch_cap() : 1;
                t = pick(true, first, count, max_take, &split);
                if (!t) continue;
            }
        }

        if (mode == PM_HIP_MODE_BATCHED && t->gpu_resident) {
            // Gather subtasks of this and other active tasks into one batch.
            BatchRun run;
            struct Claim { Task* task; ulong first, count; pmSplitInfo split; };
            std::vector<Claim> claims;
            claims.push_back({t, first, count, split});
            ulong taken = count;
            const ulong cap = batch_cap();
            {
                std::lock_guard<std::mutex> g(q_mtx);
                while (taken < cap) {
                    ulong f2 = 0, c2 = 0;
                    Task* t2 = nullptr;
                    // only tasks that can be batched (GPU-resident)
                    for (Task* x : active) {
                        if (x->gpu_resident && x->gpu_prepared && x->hip_mode == PM_HIP_MODE_BATCHED &&
                            x->has_work(true)) { t2 = x; break; }
                    }
                    if (!t2) break;
                    pmSplitInfo sp2;
                    Task* got = pick(true, f2, c2, cap - taken, &sp2);
                    if (!got) break;
                    claims.push_back({got, f2, c2, sp2});
                    taken += c2;
                }
            }
            const double s0 = now_secs();
            for (auto& cl : claims) {
                Task* tt = cl.task;
                try { ensure_task_ready(tt); } catch (Error& e) { fail_task(tt, e); }
                for (ulong sid = cl.first; sid < cl.first + cl.count; ++sid) {
                    std::unique_ptr<GpuSlot> sp(new GpuSlot());
                    sp->task = tt;
                    sp->st.reset(new Subtask());
                    sp->st->id = sid;
                    sp->st->split = cl.split;
                    sp->busy = true;
                    if (tt->status.load() == pmSuccess) {
                        try {
                            if (tt->callbacks->cb.deviceSelection && !tt->callbacks->cb.deviceSelection(tt->info(), *gpu_device))
                                raise(pmNoCompatibleDevice, "device selection rejected the HIP stub");
                            run_distribution(tt, *sp->st, gpu_device);
                            pmSubtaskInfo si = stage_resident(tt, *sp, batch_stream);
                            if (sp->st->reserved_gpu)
                                si.gpuContext.reservedGlobalMem = slot_device_alloc(*sp, sp->st->reserved_gpu);
                            upload_scratch(*sp->st, *sp, si);
                            pmStatus rs = run_gpu_callback(tt, *sp->st, si, batch, gpu_device, kBatchStatusSlot);
                            if (rs != pmSuccess) raise(rs, "GPU callback failed");
                            download_scratch(*sp->st, *sp, batch);
                        } catch (Error& e) { fail_task(tt, e); }
                    }
                    run.slots.push_back(std::move(sp));
                }
                // the stub held one pick per claim; slots release one each
            }
            const double f0 = now_secs();
            try {
                hipchk(uh_batch_flush(batch, use_graphs.load()), "batch flush");
            } catch (Error& e) {
                for (auto& cl : claims) fail_task(cl.task, e);
            }
            stat_add(&pmHipStubStats::ns_stage, (uint64_t)((f0 - s0) * 1e9));
            stat_add(&pmHipStubStats::ns_flush, (uint64_t)((now_secs() - f0) * 1e9));
            run.event = batch_events[batch_ev];
            batch_ev = (batch_ev + 1) % 4;
            hipchk(uh_event_record(run.event, batch_stream), "batch event");
            {
                std::lock_guard<std::mutex> g(stats_mtx);
                stats.batches++;
            }
            for (auto& sp : run.slots)
                for (uint i = 0; i < sp->task->spaces.size(); ++i)
                    if (sp->task->writable(i) && !sp->task->is_reduced(i)) Memory::get().mark_device_written(sp->task->spaces[i]);
            last_batch = run.slots.size();
            batches.push_back(std::move(run));
            continue;
        }

        // PLAIN / RESIDENT: one slot per subtask
        for (ulong sid = first; sid < first + count; ++sid) {
            int si_idx = -1;
            for (int k = 0; k < depth; ++k) {
                int idx = (next_slot + k) % depth;
                if (!gs.slots[idx].busy) { si_idx = idx; break; }
            }
            if (si_idx < 0) {
                // all slots busy: wait for one
                for (int k = 0; k < depth; ++k) {
                    GpuSlot& s = gs.slots[k];
                    if (s.busy) {
                        uh_event_sync(s.event);
                        const bool plain = s.plain;
                        if (!plain && !s.dev_scatters.empty()) {
                            try { scatter_resident(s.task, s, s.stream); hipchk(uh_stream_sync(s.stream), "scatter"); }
                            catch (Error& e) { fail_task(s.task, e); }
                        }
                        retire_slot(this, s, plain);
                        inflight--;
                        si_idx = k;
                        break;
                    }
                }
            }
            next_slot = (si_idx + 1) % depth;
            GpuSlot& s = gs.slots[si_idx];
            s.seq = ++slot_seq;
            std::function<bool()> backoff = [&]() {
                int oldest = -1;
                for (int k = 0; k < depth; ++k)
                    if (k != si_idx && gs.slots[k].busy && (oldest < 0 || gs.slots[k].seq < gs.slots[oldest].seq))
                        oldest = k;
                if (oldest < 0) return false;
                GpuSlot& o = gs.slots[oldest];
                uh_event_sync(o.event);
                const bool op = o.plain;
                if (!op && !o.dev_scatters.empty()) {
                    try { scatter_resident(o.task, o, o.stream); hipchk(uh_stream_sync(o.stream), "scatter"); }
                    catch (Error& e) { fail_task(o.task, e); }
                }
                retire_slot(this, o, op);
                inflight--;
                return true;
            };
            struct BackoffScope {
                explicit BackoffScope(std::function<bool()>* f) { tl_oom_backoff = f; }
                ~BackoffScope() { tl_oom_backoff = nullptr; }
            } backoff_scope(&backoff);
            s.task = t;
            s.st.reset(new Subtask());
            s.st->id = sid;
            s.st->split = split;
            s.stream = gpu_streams[si_idx];
            s.event = gpu_events[si_idx];
            s.busy = true;
            const bool plain = !t->gpu_resident || mode == PM_HIP_MODE_PLAIN;
            s.plain = plain;
            s.ma_copy = ma_copy;
            if (t->multi_assign && !ma_copy) {
                std::lock_guard<std::mutex> g(t->ma_mtx);
                if (t->gpu_first == 0) t->gpu_first = now_secs();
            }
            if (t->status.load() == pmSuccess) {
                try {
                    if (t->callbacks->cb.deviceSelection && !t->callbacks->cb.deviceSelection(t->info(), *gpu_device))
                        raise(pmNoCompatibleDevice, "device selection rejected the HIP stub");
                    if (!plain) ensure_task_ready(t);
                    run_distribution(t, *s.st, gpu_device);
                    pmSubtaskInfo info = plain ? stage_plain(t, s) : stage_resident(t, s, s.stream);
                    if (s.st->reserved_gpu) info.gpuContext.reservedGlobalMem = slot_device_alloc(s, s.st->reserved_gpu);
                    upload_scratch(*s.st, s, info);
                    pmStatus rs = run_gpu_callback(t, *s.st, info, s.stream, gpu_device, si_idx);
                    if (rs != pmSuccess) raise(rs, "GPU callback failed");
                    download_scratch(*s.st, s, s.stream);
                    if (plain) {
                        // D2H of written records into the pinned buffer (pmCudaInterface.cu:510-527)
                        uint64_t bytes = 0;
                        // sentinel compression (reference CheckSentinelCompression,
                        // pmExecutionStub.cpp:3850-3881): a built-in reduction over the task's only
                        // address space, written whole by one record, above the threshold
                        pmReductionOpType sop;
                        pmReductionDataType sdt = MAX_REDUCTION_DATA_TYPES;
                        const bool sentinel_task =
                            t->spaces.size() == 1 && t->writable(0) && t->is_reduced(0) &&
                            is_builtin_reduction(t->callbacks->cb.dataReduction, &sop, &sdt) &&
                            t->spaces[0]->length > g_sentinel_threshold.load() && s.scatters.size() == 1 &&
                            s.scatters[0].mem == 0 && s.scatters[0].src_off == 0 &&
                            s.scatters[0].len == t->spaces[0]->length;
                        if (sentinel_task) {
                            static const size_t kElem[] = {4, 4, 8, 8, 4, 8};   // pmReductionDataType sizes
                            const size_t es = kElem[sdt];
                            const size_t len = s.scatters[0].len;
                            const uint64_t elems = len / es;
                            // compressed when fewer than 40 % of the elements are non-zero
                            // (CUDA_SENTINEL_COMPRESSION_MAX_NON_SENTINELS, pmInternalDefinitions.h:224)
                            const uint64_t lim = (uint64_t)std::ceil((double)elems * 0.4);
                            const uint64_t max_nz = lim ? lim - 1 : 0;
                            const size_t cbytes = 8 + (size_t)max_nz * (es + 4);
                            if (elems * es == len && cbytes <= len) {
                                void* comp = slot_device_alloc(s, cbytes);
                                uint64_t nz = 0;
                                hipchk(uh_sentinel_compress(s.dev_scatters[0].buf, elems, (int)sdt, comp, max_nz, &nz,
                                                            s.stream), "sentinel compression");
                                if (nz <= max_nz) {
                                    const size_t got = 8 + (size_t)nz * (es + 4);
                                    hipchk(uh_memcpy_d2h(s.pinned + s.scatters[0].pinned_off, comp, got, s.stream),
                                           "D2H staging");
                                    s.scatters[0].sent = (int64_t)nz;
                                    s.scatters[0].sent_elem = (uint32_t)es;
                                    bytes += got;
                                    stat_add(&pmHipStubStats::sentinel_compressed, 1);
                                }
                            }
                        }
                        for (size_t k = 0; k < s.scatters.size();) {
                            if (s.scatters[k].sent >= 0) { ++k; continue; }
                            size_t v = k + 1, len = s.scatters[k].len;
                            while (v < s.scatters.size() &&
                                   s.scatters[v].mem == s.scatters[k].mem &&
                                   s.scatters[v].pinned_off == s.scatters[k].pinned_off + len &&
                                   (uint8_t*)s.dev_scatters[v].buf == (uint8_t*)s.dev_scatters[k].buf + len) {
                                len += s.scatters[v].len;
                                ++v;
                            }
                            hipchk(uh_memcpy_d2h(s.pinned + s.scatters[k].pinned_off, s.dev_scatters[k].buf, len,
                                                 s.stream), "D2H staging");
                            bytes += len;
                            k = v;
                        }
                        s.dev_scatters.clear();
                        stat_add(&pmHipStubStats::d2h_bytes, bytes);
                    } else {
                        for (uint i = 0; i < t->spaces.size(); ++i)
                            if (t->writable(i) && !t->is_reduced(i)) Memory::get().mark_device_written(t->spaces[i]);
                    }
                } catch (Error& e) {
                    // a failed copy is dropped: the CPU stub's run of the subtask still counts
                    if (s.ma_copy) s.copy_failed = true;
                    else fail_task(t, e);
                }
            }
            hipchk(uh_event_record(s.event, s.stream), "subtask event");
            inflight++;
        }
    }
    while (!batches.empty()) retire_batches(true);
    for (int i = 0; i < 4; ++i) uh_event_destroy(batch_events[i]);
    cache_purge(nullptr, true);
}


}  // namespace rt
}  // namespace pm
