// Internal types of the single-node pm:: runtime (libpmlib.so).
//
// Layer map (reference counterparts in parentheses):
//   api.cpp           public API, status mapping            (pmPublicDefinitions.cpp)
//   subscription.cpp  subscription records and views        (pmSubscriptionManager.cpp:397-2031)
//   memory.cpp        address spaces, HBM mirrors, budget   (pmAddressSpace, pmMemChunk, pmCache)
//   executor.cpp      tasks, per-device queues, CPU stubs,  (pmTask, pmScheduler, pmExecutionStub,
//                     the HIP stub and its three modes       pmDispatcherGPU, pmCudaInterface)
//   reduce.cpp        built-in subtask reductions           (pmReducer.cpp:593-710)
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pmPublicDefinitions.h"
#include "pmPublicUtilities.h"
#include "pmExtensions.h"

namespace pm {
namespace rt {

// Thrown inside the runtime, mapped to pmStatus at the API boundary.
struct Error {
    pmStatus status;
    std::string what;
};
[[noreturn]] void raise(pmStatus s, const std::string& what);
void set_last_error(const std::string& s);

double now_secs();

// ---------------------------------------------------------------- intervals / subscriptions

// Disjoint, non-adjacent byte intervals [offset, offset+length) keyed by offset.
class IntervalSet {
public:
    void add(size_t offset, size_t length);
    bool empty() const { return m_.empty(); }
    const std::map<size_t, size_t>& map() const { return m_; }
    size_t total() const;
private:
    std::map<size_t, size_t> m_;
};

// Subscriptions of one subtask to one address space. The subscriptions are kept as given
// (contiguous and scattered) together with their consolidated extents; the per-record union
// (what the reference materialises row by row in a map) is built only when a COMPACT view or
// record-level staging needs it — NATURAL views of HBM mirrors only use the extents.
struct MemSubscription {
    std::vector<pmSubscriptionInfo> rlin, wlin;
    std::vector<pmScatteredSubscriptionInfo> rsc, wsc;
    pmSubscriptionInfo cread{0, 0}, cwrite{0, 0};     // consolidated extents
    bool has_read = false, has_write = false;
    void add(pmSubscriptionType type, size_t offset, size_t length);
    void add_scattered(pmSubscriptionType type, const pmScatteredSubscriptionInfo& s);
    pmSubscriptionInfo unified() const;               // min offset / max span of read and write
    const IntervalSet& read() const;                  // union of read records (built lazily)
    const IntervalSet& write() const;                 // union of write records (built lazily)
private:
    mutable bool built_r_ = false, built_w_ = false;
    mutable IntervalSet read_, write_;
};

// COMPACT view geometry: union of read and write records packed back to back in address order
// (reference GetCompactedSubscription, pmSubscriptionManager.cpp:1978-2031).
struct CompactLayout {
    std::vector<size_t> src_off, dst_off, len;   // one entry per packed record
    size_t total = 0;
    size_t first_read = 0, first_write = 0;      // compact offsets of the first records
    bool has_read = false, has_write = false;
    size_t compact_offset(size_t addr) const;    // address must lie inside a record
};
CompactLayout make_compact(const MemSubscription& s);

// ---------------------------------------------------------------- address spaces

struct AddressSpace {
    size_t length = 0;
    size_t rows = 0, cols = 0;          // 2-D layout (cols in bytes), rows == 0 for linear
    uint8_t* host = nullptr;
    bool host_pinned = false;
    uint8_t* dev = nullptr;             // HBM mirror
    bool host_valid = true;
    bool dev_valid = false;
    int users = 0;                      // tasks currently holding it
    int gpu_pins = 0;                   // GPU-admitted tasks using the mirror (not evictable)
    uint64_t last_use = 0;
    uint64_t uid = 0;                   // unique id (never reused)
    std::atomic<uint64_t> version{0};   // bumped on every write; keys the PLAIN-mode cache
    std::mutex mtx;
    bool released = false;
    // Unfinished tasks using this space, in submission order, and whether each writes it
    // (guarded by Runtime::dep_mtx).
    std::vector<std::pair<struct Task*, bool>> holders;
    // multi-assigned copies still reading it (a copy that lost keeps running); release waits
    int runners = 0;
    std::condition_variable runners_cv;
};

// Bytes of the span u that lie inside the address space. The consolidated extent of a scattered
// subscription is step * count bytes from its offset (the reference's definition), which can end
// past the space when the last record is shorter than the step; copies of a whole span stop at the
// space's end.
inline size_t span_in_space(const AddressSpace* as, const pmSubscriptionInfo& u) {
    return u.offset >= as->length ? 0 : (u.length < as->length - u.offset ? u.length : as->length - u.offset);
}

class Memory {
public:
    static Memory& get();
    AddressSpace* create(size_t length, size_t rows, size_t cols);
    void release(AddressSpace* as);
    // Make the host copy current (download from HBM if the mirror is newer).
    void sync_host(AddressSpace* as);
    // copies [off, off + len) of the newest copy to dst; validity unchanged
    void read_range(AddressSpace* as, size_t off, size_t len, void* dst);
    // Make the HBM mirror current (allocate within the budget, upload if the host is newer).
    bool ensure_device(AddressSpace* as, void* stream);   // true when it enqueued an upload
    void mark_device_written(AddressSpace* as);
    void mark_host_written(AddressSpace* as);
    // HBM budget accounting for mirrors and subtask buffers.
    size_t budget() const { return budget_; }
    size_t resident() const { return resident_.load(); }
    void set_budget(size_t b) { budget_ = b; }
    size_t default_budget = 0;          // the budget chosen at initialisation
    bool try_reserve(size_t bytes);
    void unreserve(size_t bytes);
    void* device_alloc(size_t bytes);   // budgeted hipMalloc; evicts idle mirrors on pressure
    void device_free(void* p, size_t bytes);
    // Evicts mirrors that no GPU-admitted task uses (LRU) until `need` more bytes fit; the
    // spaces in `keep` are never evicted.
    void evict_idle(size_t need, const std::vector<AddressSpace*>* keep = nullptr);
    // Whether the missing mirrors of `spaces` fit the budget (after evicting idle mirrors).
    bool admit(const std::vector<AddressSpace*>& spaces);
    // Missing-mirror bytes of `spaces` (each space counted once).
    size_t missing_bytes(const std::vector<AddressSpace*>& spaces);
    // Ids of address spaces released since the last call (the HIP stub drops their cached
    // device buffers).
    std::vector<uint64_t> take_released();
    std::mutex list_mtx;
    std::vector<AddressSpace*> spaces;
    std::vector<uint64_t> released_uids;
private:
    size_t budget_ = 0;
    std::atomic<size_t> resident_{0};
    uint64_t clock_ = 0;
};

// ---------------------------------------------------------------- tasks

struct CallbackSet {
    std::string key;
    pmCallbacks cb;
};

struct Task;

struct Subtask {
    ulong id = 0;
    pmSplitInfo split;
    MemSubscription mem[MAX_MEM_SECTIONS_PER_TASK];
    size_t reserved_gpu = 0;
    bool has_launch = false;
    pmGpuLaunchConf launch;
    // scratch buffers per pmScratchBufferType (host side)
    std::map<int, std::vector<uint8_t>> scratch;
    // shadow buffers for reduced address spaces (host or device)
    void* shadow[MAX_MEM_SECTIONS_PER_TASK] = {nullptr, nullptr, nullptr};
    bool shadow_on_device[MAX_MEM_SECTIONS_PER_TASK] = {false, false, false};
    size_t shadow_len[MAX_MEM_SECTIONS_PER_TASK] = {0, 0, 0};
    pmDeviceInfo* device = nullptr;
};

enum class TaskState { Queued, Running, Done };

struct Task {
    pmTaskDetails details;
    std::vector<uint8_t> conf;                   // owned copy of taskConf
    std::vector<pmTaskMem> mem;
    std::vector<AddressSpace*> spaces;
    CallbackSet* callbacks = nullptr;
    ulong subtask_count = 0;
    std::atomic<ulong> next{0};                  // next unassigned subtask id
    std::atomic<ulong> done{0};
    std::atomic<int> status{pmSuccess};
    bool cpu_allowed = false, gpu_allowed = false;
    bool gpu_resident = false;                   // GPU subtasks use HBM mirrors (GPU-only task)
    bool gpu_prepared = false;                   // mirrors made current by the HIP stub
    int hip_mode = PM_HIP_MODE_BATCHED;          // HIP stub mode at submission
    void* dev_conf = nullptr;                    // device copy of conf for kernel-kind callbacks
    // Ordering against earlier tasks on the same address spaces (guarded by Runtime::dep_mtx):
    // the task starts when `pending` earlier conflicting tasks have completed.
    bool unordered = false;                      // submitted without waiting for earlier tasks
    int pending = 0;
    std::vector<Task*> dependents;               // later tasks waiting for this one
    // Splitting (reference pmSubtaskSplitter.cpp): on the splitting device type a subtask runs as
    // split_count pieces, each called with its pmSplitInfo, and completes with its last piece.
    int split_device = 0;                        // 0: no splitting, 1: CPU stubs, 2: the HIP stub
    uint split_count = 1;
    std::mutex split_mtx;
    std::deque<std::pair<ulong, uint>> open_splits;   // (subtask, next split id) not all handed out
    std::map<ulong, uint> split_left;                 // subtask -> pieces not yet finished
    // Whether a stub of this kind can take work from the task now (caller holds Runtime::q_mtx).
    bool has_work(bool gpu);
    // Multi-assign (reference pmExecutionStub.cpp:1058-1170, at most one extra copy as
    // MAX_SUBTASK_MULTI_ASSIGN_COUNT 2): the idle HIP stub may run a copy of a subtask that a CPU
    // stub is executing; the copy that finishes first commits its output and completes the
    // subtask, the other's output is dropped. CPU subtasks of such a task write into private
    // shadows that are copied to the address space on commit.
    bool multi_assign = false;
    struct CpuRun {
        double start = 0;
        bool duplicated = false;                 // a HIP copy was started
        bool committed = false;
    };
    std::mutex ma_mtx;
    std::map<ulong, CpuRun> cpu_runs;            // subtasks executing on CPU stubs
    double cpu_secs = 0;                         // summed CPU subtask durations
    ulong cpu_finished = 0;
    double gpu_first = 0, gpu_last = 0;          // first HIP staging, last HIP retirement
    ulong gpu_finished = 0;
    // copies still executing (CPU subtasks, losers included, and HIP copies); the task object
    // and its address spaces stay alive until they end (guarded by mtx, signalled on cv)
    int runners = 0;
    // First finisher of subtask `sid` takes the commit; false when the other copy already has.
    bool claim_commit(ulong sid, bool gpu_copy);
    double submit_time = 0, start_time = 0, end_time = 0;
    double deadline = 0;
    std::mutex mtx;
    std::condition_variable cv;
    TaskState state = TaskState::Queued;
    std::atomic<bool> finished{false};  // state == Done, readable without the mutex
    uint64_t seq = 0;
    // reduction: completed subtasks awaiting a partner
    std::unique_ptr<Subtask> pending_reduce;
    std::mutex reduce_mtx;
    std::vector<std::unique_ptr<Subtask>> reduce_queue;
    // Writes to a READ_WRITE space whose subtasks may read what other subtasks write
    // (pmTaskMem::disjointReadWritesAcrossSubtasks false, the default) are held back until the
    // task ends, so that every subtask reads the task's input (reference
    // pmTask::DoSubtasksNeedShadowMemory, pmTask.cpp:749-754, committed as pmTask.cpp:1173-1191).
    bool deferred(uint i) const;
    bool has_deferred = false;
    struct DeferredWrite {
        ulong subtask;
        uint mem;
        size_t offset;
        std::vector<uint8_t> bytes;
    };
    std::mutex deferred_mtx;
    std::vector<DeferredWrite> deferred_writes;
    void defer_write(ulong subtask, uint mem, size_t offset, const void* src, size_t len);
    // Data redistribution (reference pmRedistributor.cpp, pmSubscriptionManager.cpp:861-877): the
    // dataRedistribution callback of every completed subtask registers ranges of the subtask's
    // output (offsets relative to its consolidated write subscription) with an order number; at
    // the end of the task the address space is rewritten as the concatenation of all ranges in
    // ascending order (within an order by subtask, then in registration order), and the bytes
    // after them are zeroed.
    struct RedistRecord {
        uint order;
        ulong subtask;
        uint seq;
        size_t offset, length;
    };
    std::mutex redist_mtx;
    std::map<uint, std::vector<RedistRecord>> redist;                  // per task memory index
    std::map<uint, std::vector<pmRedistributionMetadata>> redist_meta; // (order, bytes) after the task
    pmTaskInfo info();
    bool is_reduced(uint i) const;
    bool writable(uint i) const;
    bool read_only(uint i) const;
    bool write_only(uint i) const;
    pmSubscriptionVisibilityType visibility(uint i, bool gpu) const;
};

// Thread-local context of the data-distribution callback being run, so that
// pmSubscribeToMemory & co. can find the subtask they describe.
struct DistributionContext {
    Task* task = nullptr;
    Subtask* subtask = nullptr;
    pmDeviceHandle device = nullptr;
    bool distribution = true;    // false: a subtask callback (scratch buffers only, no subscribing)
    bool redistribution = false; // a dataRedistribution callback (pmRedistributeData allowed)
};
DistributionContext*& current_context();

// Makes the subtask's scratch buffers reachable (pmGetScratchBuffer) while one of its subtask
// callbacks runs on this thread, as the reference allows (matrixTranspose.cpp:325 takes its
// PRE_SUBTASK_TO_SUBTASK buffer inside subtask_cpu).
struct SubtaskScope {
    DistributionContext ctx;
    DistributionContext* saved;
    SubtaskScope(Task* t, Subtask* st, pmDeviceHandle dev) {
        ctx.task = t;
        ctx.subtask = st;
        ctx.device = dev;
        ctx.distribution = false;
        saved = current_context();
        current_context() = &ctx;
    }
    ~SubtaskScope() { current_context() = saved; }
    SubtaskScope(const SubtaskScope&) = delete;
    SubtaskScope& operator=(const SubtaskScope&) = delete;
};

// Status word of kernel-kind callbacks launched from a batch (slots 0..15 are the per-subtask
// slots of the PLAIN / RESIDENT modes).
constexpr int kBatchStatusSlot = 16;

class Runtime {
public:
    static Runtime* get();               // null before pmInitialize
    static void init();
    static void finalize();
    std::vector<pmDeviceInfo*> cpu_devices;
    pmDeviceInfo* gpu_device = nullptr;
    int gpu_count = 0;
    // callbacks
    std::mutex cb_mtx;
    std::map<std::string, std::unique_ptr<CallbackSet>> callbacks;
    // task queue
    void submit(Task* t);
    void wait(Task* t);
    std::mutex q_mtx;
    std::condition_variable q_cv;
    std::vector<Task*> active;           // queued or running, ordered by (priority, seq)
    uint64_t seq = 0;
    bool stopping = false;
    std::atomic<uint64_t> submitted{0};   // tasks submitted so far (the idle GPU thread spins on it)
    // stubs
    std::vector<std::thread> cpu_threads;
    std::thread gpu_thread;
    void cpu_loop(int idx);
    void gpu_loop();
    // mode of the HIP stub
    std::atomic<int> hip_mode{PM_HIP_MODE_BATCHED};
    std::atomic<int> hip_view{SUBSCRIPTION_NATURAL};
    std::atomic<int> use_graphs{1};
    std::atomic<int> pipeline_depth{4};
    std::atomic<int> batch_limit{4096};
    std::atomic<int> cpu_workers{0};
    std::atomic<int> split_factor_cpu{0};    // pieces per split CPU subtask (0: one per CPU stub)
    std::atomic<int> split_factor_gpu{0};    // pieces per split GPU subtask (0: one per GPU stub)
    pmHipStubStats stats{};
    std::mutex stats_mtx;
    void* gpu_streams[16] = {};
    void* gpu_events[16] = {};
    void* batch = nullptr;
    void* batch_stream = nullptr;
    void* copy_stream = nullptr;
    void* dev_info = nullptr;            // device copy of gpu_device (kernel-kind callbacks)
    void* dev_status = nullptr;          // kBatchStatusSlot + 1 pmStatus words, written by kernels
    std::atomic<int> gpu_admitted{0};    // GPU-resident tasks admitted and not yet complete
    // ordering of tasks that share address spaces
    std::mutex dep_mtx;
    std::vector<Task*> deferred_ready;   // released while q_mtx was held; made ready by drain_ready
    void make_ready(Task* t);            // host copies made current, then queued for the stubs
    void drain_ready();                  // caller does not hold q_mtx
    // common helpers
    // Takes work for a stub: subtasks [first, first + count), or one piece of a split subtask
    // (count 1, *split set; split->splitCount stays 0 otherwise).
    Task* pick(bool gpu, ulong& first, ulong& count, ulong max_count, pmSplitInfo* split = nullptr);
    void finish_subtasks(Task* t, ulong n);
    // A subtask or one piece of it finished.
    void finish_piece(Task* t, ulong subtask, const pmSplitInfo& split);
    // Drops the pieces of open split records that were never handed out (cancellation).
    void drop_open_splits(Task* t);
    void complete(Task* t);
    // multi-assign: tasks whose CPU subtasks the HIP stub may copy
    std::mutex ma_list_mtx;
    std::vector<Task*> ma_tasks;
    Task* ma_candidate(ulong& sid, double& retry_in);
    void runner_enter(Task* t);
    void runner_exit(Task* t);
    // Cancels the unpicked subtasks of a queued task with `status` (caller holds q_mtx).
    void cancel_locked(Task* t, pmStatus status);
private:
    static Runtime* inst_;
};

// Runs the dataDistribution callback of one subtask and fills its subscriptions.
void run_distribution(Task* t, Subtask& st, pmDeviceInfo* dev);
// Runs the dataRedistribution callback of a completed subtask (its output committed).
void run_redistribution(Task* t, Subtask& st, pmDeviceInfo* dev);
// Rewrites the task's redistributed address spaces (at task completion).
void commit_redistribution(Task* t);
// Builds the pmSubtaskInfo of a CPU (host, NATURAL/COMPACT) subtask; may allocate shadows.
// private_writes: writable NATURAL views point into shadows (multi-assigned CPU subtasks).
pmSubtaskInfo host_subtask_info(Task* t, Subtask& st, std::vector<std::vector<uint8_t>>& compact_bufs,
                                bool private_writes = false);
// Size above which PLAIN-mode reduction outputs go back sentinel-compressed
// (pmSetSentinelCompressionThreshold; the reference's 1 GiB by default).
extern std::atomic<size_t> g_sentinel_threshold;

// Reduction helpers
bool is_builtin_reduction(pmDataReductionCallback cb, pmReductionOpType* op, pmReductionDataType* dt);
void reduce_host(void* dst, const void* src, size_t bytes, pmReductionOpType op, pmReductionDataType dt);

}  // namespace rt
}  // namespace pm
