// Subscription records and subtask views.
//
// Follows the reference's subscription manager semantics:
//   - a READ_WRITE subscription registers both a read and a write record
//     (pmSubscriptionManager.cpp:425-433);
//   - records of one kind are kept as a union of byte ranges; overlapping and adjacent ranges
//     merge (AddSubscriptionRecordToMap, pmSubscriptionManager.cpp:744-843);
//   - the consolidated extent of a kind is the min offset / max span of its subscriptions
//     (pmSubscriptionManager.cpp:475-489, 508-523); a scattered subscription spans
//     step * count bytes from its offset;
//   - the NATURAL view of a read-only space starts at the consolidated read extent, that of a
//     writable space at the unified read/write extent (GetSubtaskInfo, 1882-1976);
//   - the COMPACT view packs the union of read and write records in address order
//     (GetCompactedSubscription, 1978-2031); readPtr / writePtr point at the packed position of
//     the first read / write record.
#include "pm_internal.h"
#include "unicorn_hip.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <memory>

namespace pm {
namespace rt {

void IntervalSet::add(size_t offset, size_t length) {
    if (!length) return;
    size_t lo = offset, hi = offset + length;
    // records usually arrive in ascending address order: append or extend the last interval
    if (!m_.empty()) {
        auto last = std::prev(m_.end());
        const size_t lend = last->first + last->second;
        if (lo > lend) { m_.emplace_hint(m_.end(), lo, hi - lo); return; }
        if (lo >= last->first) { if (hi > lend) last->second = hi - last->first; return; }
    }
    // first interval that could overlap or touch [lo, hi): the last one starting <= lo
    auto it = m_.upper_bound(lo);
    if (it != m_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second >= lo) it = prev;
    }
    while (it != m_.end() && it->first <= hi) {
        lo = std::min(lo, it->first);
        hi = std::max(hi, it->first + it->second);
        it = m_.erase(it);
    }
    m_[lo] = hi - lo;
}

size_t IntervalSet::total() const {
    size_t t = 0;
    for (auto& kv : m_) t += kv.second;
    return t;
}

static void extend(pmSubscriptionInfo& c, bool& has, size_t off, size_t span) {
    if (!has) {
        c = pmSubscriptionInfo(off, span);
        has = true;
        return;
    }
    size_t lo = std::min(c.offset, off);
    size_t hi = std::max(c.offset + c.length, off + span);
    c = pmSubscriptionInfo(lo, hi - lo);
}

void MemSubscription::add(pmSubscriptionType type, size_t offset, size_t length) {
    if (!length) return;
    if (type == READ_SUBSCRIPTION || type == READ_WRITE_SUBSCRIPTION) {
        rlin.emplace_back(offset, length);
        extend(cread, has_read, offset, length);
        built_r_ = false;
    }
    if (type == WRITE_SUBSCRIPTION || type == READ_WRITE_SUBSCRIPTION) {
        wlin.emplace_back(offset, length);
        extend(cwrite, has_write, offset, length);
        built_w_ = false;
    }
}

void MemSubscription::add_scattered(pmSubscriptionType type, const pmScatteredSubscriptionInfo& s) {
    if (!s.size || !s.count) return;
    const bool r = (type == READ_SUBSCRIPTION || type == READ_WRITE_SUBSCRIPTION);
    const bool w = (type == WRITE_SUBSCRIPTION || type == READ_WRITE_SUBSCRIPTION);
    // consolidated extent of a scattered subscription: step * count bytes from its offset
    // (reference pmSubscriptionManager.cpp:508-523)
    if (r) { rsc.push_back(s); extend(cread, has_read, s.offset, s.step * s.count); built_r_ = false; }
    if (w) { wsc.push_back(s); extend(cwrite, has_write, s.offset, s.step * s.count); built_w_ = false; }
}

static void build_union(IntervalSet& u, const std::vector<pmSubscriptionInfo>& lin,
                        const std::vector<pmScatteredSubscriptionInfo>& sc) {
    u = IntervalSet();
    for (const auto& l : lin) u.add(l.offset, l.length);
    for (const auto& s : sc) {
        if (s.size == s.step) { u.add(s.offset, s.size * s.count); continue; }   // adjacent records
        for (size_t i = 0; i < s.count; ++i) u.add(s.offset + i * s.step, s.size);
    }
}

const IntervalSet& MemSubscription::read() const {
    if (!built_r_) { build_union(read_, rlin, rsc); built_r_ = true; }
    return read_;
}

const IntervalSet& MemSubscription::write() const {
    if (!built_w_) { build_union(write_, wlin, wsc); built_w_ = true; }
    return write_;
}

pmSubscriptionInfo MemSubscription::unified() const {
    if (!has_read && !has_write) return pmSubscriptionInfo(0, 0);
    if (has_read && !has_write) return cread;
    if (!has_read && has_write) return cwrite;
    size_t lo = std::min(cread.offset, cwrite.offset);
    size_t hi = std::max(cread.offset + cread.length, cwrite.offset + cwrite.length);
    return pmSubscriptionInfo(lo, hi - lo);
}

size_t CompactLayout::compact_offset(size_t addr) const {
    // records are sorted by src_off
    auto it = std::upper_bound(src_off.begin(), src_off.end(), addr);
    if (it == src_off.begin()) raise(pmFatalError, "compact_offset: address before first record");
    size_t i = (size_t)(it - src_off.begin()) - 1;
    if (addr >= src_off[i] + len[i]) raise(pmFatalError, "compact_offset: address outside records");
    return dst_off[i] + (addr - src_off[i]);
}

CompactLayout make_compact(const MemSubscription& s) {
    IntervalSet u = s.read();
    for (auto& kv : s.write().map()) u.add(kv.first, kv.second);
    CompactLayout c;
    size_t pos = 0;
    for (auto& kv : u.map()) {
        c.src_off.push_back(kv.first);
        c.dst_off.push_back(pos);
        c.len.push_back(kv.second);
        pos += kv.second;
    }
    c.total = pos;
    if (!s.read().empty()) {
        c.has_read = true;
        c.first_read = c.compact_offset(s.read().map().begin()->first);
    }
    if (!s.write().empty()) {
        c.has_write = true;
        c.first_write = c.compact_offset(s.write().map().begin()->first);
    }
    return c;
}

DistributionContext*& current_context() {
    static thread_local DistributionContext* ctx = nullptr;
    return ctx;
}

void run_distribution(Task* t, Subtask& st, pmDeviceInfo* dev) {
    st.device = dev;
    pmDataDistributionCallback dd = t->callbacks->cb.dataDistribution;
    if (!dd) return;
    DistributionContext ctx;
    ctx.task = t;
    ctx.subtask = &st;
    ctx.device = dev->deviceHandle;
    DistributionContext* saved = current_context();
    current_context() = &ctx;
    pmSubtaskInfo si;
    si.subtaskId = st.id;
    si.memCount = (uint)t->spaces.size();
    si.splitInfo = st.split;
    pmStatus s = dd(t->info(), *dev, si);
    current_context() = saved;
    if (s != pmSuccess) raise(s, "dataDistribution callback failed");
}

void run_redistribution(Task* t, Subtask& st, pmDeviceInfo* dev) {
    pmDataRedistributionCallback cb = t->callbacks->cb.dataRedistribution;
    if (!cb) return;
    DistributionContext ctx;
    ctx.task = t;
    ctx.subtask = &st;
    ctx.device = dev->deviceHandle;
    ctx.distribution = false;
    ctx.redistribution = true;
    DistributionContext* saved = current_context();
    current_context() = &ctx;
    // the subtask's views of the committed host copy; a task that ran resident on the HIP stub
    // keeps its bytes in HBM, and its callback sees the lengths only (the reference's
    // redistribution callbacks read their scratch buffers, not the data)
    pmSubtaskInfo si;
    si.subtaskId = st.id;
    si.memCount = (uint)t->spaces.size();
    si.splitInfo = st.split;
    for (uint i = 0; i < t->spaces.size(); ++i) {
        pmSubscriptionInfo u = st.mem[i].unified();
        if (!u.length) continue;
        pmMemInfo& mi = si.memInfo[i];
        mi.length = u.length;
        mi.visibilityType = SUBSCRIPTION_NATURAL;
        if (t->gpu_resident) continue;
        mi.ptr = t->spaces[i]->host + u.offset;
        if (st.mem[i].has_read) mi.readPtr = t->spaces[i]->host + st.mem[i].cread.offset;
        if (st.mem[i].has_write) mi.writePtr = t->spaces[i]->host + st.mem[i].cwrite.offset;
    }
    pmStatus s = cb(t->info(), *dev, si);
    current_context() = saved;
    if (s != pmSuccess) raise(s, "dataRedistribution callback failed");
}

// Runs body(0 .. n-1) on up to `threads` threads (the calling one included), in contiguous ranges.
template <typename F>
static void parallel_for(size_t n, unsigned threads, F body) {
    if (threads <= 1 || n <= 1) {
        for (size_t k = 0; k < n; ++k) body(k);
        return;
    }
    threads = (unsigned)std::min<size_t>(threads, n);
    std::vector<std::thread> th;
    const size_t per = (n + threads - 1) / threads;
    for (unsigned j = 1; j < threads; ++j)
        th.emplace_back([&, j] {
            for (size_t k = j * per; k < std::min(n, (j + 1) * per); ++k) body(k);
        });
    for (size_t k = 0; k < std::min(n, per); ++k) body(k);
    for (auto& x : th) x.join();
}

// The same concatenation when the address space's current copy is its HBM mirror: the records
// are copied by one kernel into a zeroed device buffer, which then replaces the mirror's bytes;
// the host copy stays stale until something reads it.
static void redistribute_on_device(AddressSpace* as, const std::vector<Task::RedistRecord>& recs,
                                   const std::vector<size_t>& dst, size_t total) {
    const size_t n = recs.size();
    const size_t words = (as->length + 3) / 4;
    std::vector<uint64_t> tab(3 * n);
    uint64_t maxlen = 0;
    for (size_t k = 0; k < n; ++k) {
        tab[k] = recs[k].offset;
        tab[n + k] = dst[k];
        tab[2 * n + k] = recs[k].length;
        maxlen = std::max<uint64_t>(maxlen, recs[k].length);
    }
    Memory& M = Memory::get();
    void* tmp = M.device_alloc(words * 4);
    void* dtab = n ? M.device_alloc(tab.size() * 8) : nullptr;
    void* stream = nullptr;
    int rc = uh_stream_create(&stream);
    if (!rc && n) rc = uh_memcpy_h2d(dtab, tab.data(), tab.size() * 8, stream);
    if (!rc) rc = uh_fill_u32(tmp, 0, words, stream);
    if (!rc && n)
        rc = uh_copy_records(tmp, as->dev, (const uint64_t*)dtab, (const uint64_t*)dtab + n, (const uint64_t*)dtab + 2 * n,
                             (uint32_t)n, maxlen, stream);
    if (!rc) rc = uh_memcpy_d2d(as->dev, tmp, as->length, stream);
    if (!rc) rc = uh_stream_sync(stream);
    const std::string err = rc ? std::string(uh_last_error()) : std::string();
    if (stream) uh_stream_destroy(stream);
    M.device_free(tmp, words * 4);
    if (dtab) M.device_free(dtab, tab.size() * 8);
    if (rc) raise(pmGraphicsCardError, "redistribution in HBM: " + err);
    (void)total;
    M.mark_device_written(as);
}

void commit_redistribution(Task* t) {
    std::lock_guard<std::mutex> g(t->redist_mtx);
    for (auto& kv : t->redist) {
        const uint i = kv.first;
        auto& recs = kv.second;
        std::stable_sort(recs.begin(), recs.end(), [](const Task::RedistRecord& a, const Task::RedistRecord& b) {
            if (a.order != b.order) return a.order < b.order;
            if (a.subtask != b.subtask) return a.subtask < b.subtask;
            return a.seq < b.seq;
        });
        AddressSpace* as = t->spaces[i];
        // destination of every record (its position in the concatenation) and the metadata
        std::vector<size_t> dst(recs.size());
        size_t pos = 0;
        std::vector<pmRedistributionMetadata>& meta = t->redist_meta[i];
        meta.clear();
        for (size_t k = 0; k < recs.size(); ++k) {
            const auto& r = recs[k];
            if (pos + r.length > as->length) raise(pmInvalidOffset, "redistributed ranges exceed the address space");
            dst[k] = pos;
            pos += r.length;
            if (meta.empty() || meta.back().order != r.order) meta.push_back(pmRedistributionMetadata(r.order, 0));
            meta.back().count += (uint)r.length;
        }
        if (as->dev && as->dev_valid && !as->host_valid) {
            redistribute_on_device(as, recs, dst, pos);
            continue;
        }
        // gather into a buffer of the concatenation's size, then write it back and clear the rest
        // (the reference builds a fresh address space of the same length); large spaces are
        // copied by several threads
        std::unique_ptr<uint8_t[]> out(new (std::nothrow) uint8_t[pos ? pos : 1]);
        if (!out) raise(pmMemoryError, "redistribution buffer");
        const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>(16, (as->length >> 26) + 1));
        parallel_for(recs.size(), nt, [&](size_t k) { memcpy(out.get() + dst[k], as->host + recs[k].offset, recs[k].length); });
        const size_t piece = ((as->length + nt - 1) / nt + 4095) & ~(size_t)4095;
        parallel_for(nt, nt, [&](size_t j) {
            const size_t a = j * piece, b = std::min(as->length, a + piece);
            if (a >= b) return;
            if (a < pos) memcpy(as->host + a, out.get() + a, std::min(b, pos) - a);
            if (b > pos) memset(as->host + std::max(a, pos), 0, b - std::max(a, pos));
        });
        Memory::get().mark_host_written(as);
    }
}

// Host view of a subtask for the CPU stubs. NATURAL views point into the address space (or a
// shadow buffer for reduced spaces); COMPACT views are packed into compact_bufs.
pmSubtaskInfo host_subtask_info(Task* t, Subtask& st, std::vector<std::vector<uint8_t>>& compact_bufs,
                                bool private_writes) {
    pmSubtaskInfo si;
    si.subtaskId = st.id;
    si.memCount = (uint)t->spaces.size();
    si.splitInfo = st.split;
    compact_bufs.resize(t->spaces.size());
    for (uint i = 0; i < t->spaces.size(); ++i) {
        AddressSpace* as = t->spaces[i];
        MemSubscription& ms = st.mem[i];
        pmMemInfo mi;
        mi.visibilityType = t->visibility(i, false);
        if (mi.visibilityType == SUBSCRIPTION_NATURAL) {
            if (t->read_only(i)) {
                if (ms.has_read) {
                    mi.ptr = mi.readPtr = as->host + ms.cread.offset;
                    mi.length = ms.cread.length;
                }
            } else {
                pmSubscriptionInfo u = ms.unified();
                if (u.length) {
                    uint8_t* base;
                    if (t->is_reduced(i) || private_writes || t->deferred(i)) {
                        if (!st.shadow[i]) {
                            st.shadow[i] = malloc(u.length);
                            if (!st.shadow[i]) raise(pmMemoryError, "shadow allocation");
                            st.shadow_len[i] = u.length;
                            st.shadow_on_device[i] = false;
                            const size_t n = span_in_space(as, u);
                            memcpy(st.shadow[i], as->host + u.offset, n);
                            if (n < u.length) memset((uint8_t*)st.shadow[i] + n, 0, u.length - n);
                        }
                        base = (uint8_t*)st.shadow[i];
                    } else {
                        base = as->host + u.offset;
                    }
                    mi.ptr = base;
                    mi.length = u.length;
                    if (t->write_only(i)) {
                        mi.writePtr = mi.ptr;
                    } else {
                        mi.readPtr = ms.has_read ? base + (ms.cread.offset - u.offset) : nullptr;
                        mi.writePtr = ms.has_write ? base + (ms.cwrite.offset - u.offset) : nullptr;
                    }
                }
            }
        } else {
            CompactLayout cl = make_compact(ms);
            std::vector<uint8_t>& buf = compact_bufs[i];
            buf.assign(cl.total, 0);
            for (size_t r = 0; r < cl.src_off.size(); ++r)
                memcpy(buf.data() + cl.dst_off[r], as->host + cl.src_off[r], cl.len[r]);
            mi.ptr = cl.total ? buf.data() : nullptr;
            mi.length = cl.total;
            if (cl.has_read) mi.readPtr = buf.data() + cl.first_read;
            if (cl.has_write) mi.writePtr = buf.data() + cl.first_write;
        }
        si.memInfo[i] = mi;
    }
    return si;
}

}  // namespace rt
}  // namespace pm
