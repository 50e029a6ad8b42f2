// Address spaces and their HBM mirrors.
//
// The reference keeps address spaces in host memory and treats device memory as a cache of
// exact subscriptions, purged whenever a space is written (pmExecutionStub.cpp:2916-3194,
// pmAddressSpace.cpp:591-594). Here an address space may own a whole-space HBM mirror that
// stays valid across tasks; host and mirror each carry a validity flag and the runtime copies
// whole spaces between them only when the side about to be used is stale. The sum of mirrors
// and subtask buffers is held under an HBM budget; idle mirrors are evicted LRU (written back
// to the host first when they hold the only current copy).
#include "pm_internal.h"
#include "unicorn_hip.h"

#include <algorithm>
#include <cstring>
#include <cstdlib>

namespace pm {
namespace rt {

Memory& Memory::get() {
    static Memory m;
    return m;
}

AddressSpace* Memory::create(size_t length, size_t rows, size_t cols) {
    static std::atomic<uint64_t> next_uid{1};
    AddressSpace* as = new AddressSpace();
    as->uid = next_uid++;
    as->length = length;
    as->rows = rows;
    as->cols = cols;
    Runtime* rtm = Runtime::get();
    if (rtm && rtm->gpu_count > 0) {
        void* p = nullptr;
        if (uh_host_alloc(&p, length ? length : 1) == 0 && p) {
            as->host = (uint8_t*)p;
            as->host_pinned = true;
        }
    }
    if (!as->host) {
        void* p = nullptr;
        if (posix_memalign(&p, 4096, length ? length : 1) != 0) {
            delete as;
            raise(pmMemoryError, "host allocation of address space failed");
        }
        as->host = (uint8_t*)p;
    }
    std::lock_guard<std::mutex> g(list_mtx);
    spaces.push_back(as);
    return as;
}

void Memory::release(AddressSpace* as) {
    {
        // a multi-assigned copy that lost may still be reading it
        std::unique_lock<std::mutex> g(as->mtx);
        as->runners_cv.wait(g, [&] { return as->runners == 0; });
    }
    {
        std::lock_guard<std::mutex> g(list_mtx);
        for (auto it = spaces.begin(); it != spaces.end(); ++it)
            if (*it == as) { spaces.erase(it); break; }
        Runtime* r = Runtime::get();
        if (r && r->gpu_count > 0) released_uids.push_back(as->uid);
    }
    std::lock_guard<std::mutex> g(as->mtx);
    if (as->dev) {
        uh_device_sync();
        device_free(as->dev, as->length);
        as->dev = nullptr;
    }
    if (as->host_pinned) uh_host_free(as->host);
    else free(as->host);
    as->host = nullptr;
    delete as;
}

bool Memory::try_reserve(size_t bytes) {
    size_t cur = resident_.load();
    while (true) {
        if (cur + bytes > budget_) return false;
        if (resident_.compare_exchange_weak(cur, cur + bytes)) return true;
    }
}

void Memory::unreserve(size_t bytes) { resident_ -= bytes; }

void Memory::evict_idle(size_t need, const std::vector<AddressSpace*>* keep) {
    // Evict least-recently-used mirrors that no GPU-admitted task uses. A queued task that is
    // not admitted yet re-uploads its spaces on admission; the host copy is made current first.
    auto kept = [&](AddressSpace* as) {
        return keep && std::find(keep->begin(), keep->end(), as) != keep->end();
    };
    while (resident_.load() + need > budget_) {
        AddressSpace* victim = nullptr;
        {
            std::lock_guard<std::mutex> g(list_mtx);
            for (AddressSpace* as : spaces) {
                if (!as->dev || as->gpu_pins > 0 || kept(as)) continue;
                if (!victim || as->last_use < victim->last_use) victim = as;
            }
        }
        if (!victim) return;
        std::lock_guard<std::mutex> g(victim->mtx);
        if (!victim->dev || victim->gpu_pins > 0) continue;
        if (victim->dev_valid && !victim->host_valid) {
            uh_memcpy_d2h(victim->host, victim->dev, victim->length, nullptr);
            uh_stream_sync(nullptr);
            victim->host_valid = true;
        }
        device_free(victim->dev, victim->length);
        victim->dev = nullptr;
        victim->dev_valid = false;
        Runtime* r = Runtime::get();
        if (r) { std::lock_guard<std::mutex> s(r->stats_mtx); r->stats.evictions++; }
    }
}

std::vector<uint64_t> Memory::take_released() {
    std::lock_guard<std::mutex> g(list_mtx);
    std::vector<uint64_t> out;
    out.swap(released_uids);
    return out;
}

size_t Memory::missing_bytes(const std::vector<AddressSpace*>& req) {
    size_t need = 0;
    for (size_t i = 0; i < req.size(); ++i) {
        if (std::find(req.begin(), req.begin() + i, req[i]) != req.begin() + i) continue;   // listed twice
        std::lock_guard<std::mutex> g(req[i]->mtx);
        if (!req[i]->dev) need += req[i]->length ? req[i]->length : 1;
    }
    return need;
}

bool Memory::admit(const std::vector<AddressSpace*>& req) {
    const size_t need = missing_bytes(req);
    if (resident_.load() + need <= budget_) return true;
    evict_idle(need, &req);
    return resident_.load() + need <= budget_;
}

void* Memory::device_alloc(size_t bytes) {
    if (!bytes) return nullptr;
    if (!try_reserve(bytes)) {
        evict_idle(bytes);
        if (!try_reserve(bytes)) raise(pmMemoryError, "HBM budget exhausted");
    }
    void* p = nullptr;
    if (uh_malloc(&p, bytes) != 0 || !p) {
        unreserve(bytes);
        raise(pmMemoryError, std::string("hipMalloc failed: ") + uh_last_error());
    }
    return p;
}

void Memory::device_free(void* p, size_t bytes) {
    if (!p) return;
    uh_free(p);
    unreserve(bytes);
}

void Memory::sync_host(AddressSpace* as) {
    std::lock_guard<std::mutex> g(as->mtx);
    if (as->host_valid) return;
    if (as->dev && as->dev_valid) {
        if (uh_memcpy_d2h(as->host, as->dev, as->length, nullptr) || uh_stream_sync(nullptr))
            raise(pmGraphicsCardError, std::string("download of address space: ") + uh_last_error());
        Runtime* r = Runtime::get();
        if (r) { std::lock_guard<std::mutex> s(r->stats_mtx); r->stats.d2h_bytes += as->length; }
    }
    as->host_valid = true;
}

void Memory::read_range(AddressSpace* as, size_t off, size_t len, void* dst) {
    std::lock_guard<std::mutex> g(as->mtx);
    if (as->host_valid || !as->dev || !as->dev_valid) {
        memcpy(dst, as->host + off, len);
        return;
    }
    if (uh_memcpy_d2h(dst, as->dev + off, len, nullptr) || uh_stream_sync(nullptr))
        raise(pmGraphicsCardError, std::string("read of address space range: ") + uh_last_error());
    Runtime* r = Runtime::get();
    if (r) { std::lock_guard<std::mutex> s(r->stats_mtx); r->stats.d2h_bytes += len; }
}

bool Memory::ensure_device(AddressSpace* as, void* stream) {
    std::lock_guard<std::mutex> g(as->mtx);
    as->last_use = ++clock_;
    bool copied = false;
    if (!as->dev) {
        as->dev = (uint8_t*)device_alloc(as->length ? as->length : 1);
        as->dev_valid = false;
    }
    if (!as->dev_valid) {
        if (as->host_valid && as->length) {
            if (uh_memcpy_h2d(as->dev, as->host, as->length, stream))
                raise(pmGraphicsCardError, std::string("upload of address space: ") + uh_last_error());
            copied = true;
            Runtime* r = Runtime::get();
            if (r) { std::lock_guard<std::mutex> s(r->stats_mtx); r->stats.h2d_bytes += as->length; }
        }
        as->dev_valid = true;
    }
    return copied;
}

void Memory::mark_device_written(AddressSpace* as) {
    std::lock_guard<std::mutex> g(as->mtx);
    as->version++;
    as->dev_valid = true;
    as->host_valid = false;
}

void Memory::mark_host_written(AddressSpace* as) {
    std::lock_guard<std::mutex> g(as->mtx);
    as->version++;
    as->host_valid = true;
    as->dev_valid = false;
}

}  // namespace rt
}  // namespace pm
