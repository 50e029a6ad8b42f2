// Serving driver: a catalogue of resident datasets and a timed stream of requests over them.
//
// There is no reference counterpart (the reference's tasks are submitted by the app itself);
// this is the request loop of the served mode. A request names a dataset and an arrival time. It
// runs one task of that dataset's app on a free clone of the dataset (same inputs, its own
// outputs), then brings the output to the host. A pageRank request is one init task plus
// `iterations` iteration tasks.
//
// `slots` worker threads serve the stream in arrival order. Each waits for its request's arrival
// time, so a request that arrives while every slot is busy queues, and its latency includes the
// wait. The runtime underneath admits the tasks against the HBM budget, keeps address spaces
// resident, and, in the BATCHED stub mode, fuses subtasks of all tasks on the device into
// shared persistent-kernel launches (continuous batching). In the PLAIN mode every subtask is a
// separate launch with its own staging.
#include "app_common.h"

#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

extern "C" {
void* ua_mm_clone(void* h);
int ua_mm_fetch(void* h, uint64_t* checksum);
void* ua_if_clone(void* h);
int ua_if_fetch(void* h, uint64_t* checksum);
void* ua_pr_clone(void* h);
int ua_pr_fetch(void* h, uint64_t* checksum);
}

namespace serve {

enum Kind { KIND_MM = 0, KIND_IF = 1, KIND_PR = 2 };

// xoshiro256** (Blackman and Vigna), seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (auto& v : s) {
            seed += 0x9e3779b97f4a7c15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            v = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Dataset {
    Kind kind;
    void* base = nullptr;               // the instance that owns the inputs
    int iterations = 0;                 // pageRank
    std::mutex mtx;
    std::vector<void*> free_clones;     // clones not serving a request
    std::vector<void*> all_clones;
};

struct Server {
    std::vector<Dataset*> ds;
};

static void* clone_of(Dataset* d) {
    switch (d->kind) {
        case KIND_MM: return ua_mm_clone(d->base);
        case KIND_IF: return ua_if_clone(d->base);
        default: return ua_pr_clone(d->base);
    }
}

static void* acquire(Dataset* d) {
    {
        std::lock_guard<std::mutex> g(d->mtx);
        if (!d->free_clones.empty()) {
            void* c = d->free_clones.back();
            d->free_clones.pop_back();
            return c;
        }
    }
    void* c = clone_of(d);
    if (c) {
        std::lock_guard<std::mutex> g(d->mtx);
        d->all_clones.push_back(c);
    }
    return c;
}

static void release(Dataset* d, void* c) {
    std::lock_guard<std::mutex> g(d->mtx);
    d->free_clones.push_back(c);
}

static int run_request(Dataset* d, int device, int fetch, uint64_t* checksum) {
    void* c = acquire(d);
    if (!c) return -1;
    int rc;
    switch (d->kind) {
        case KIND_MM:
            rc = ua_mm_run(c, device, nullptr);
            if (!rc && fetch) rc = ua_mm_fetch(c, checksum);
            break;
        case KIND_IF:
            rc = ua_if_run(c, device, nullptr);
            if (!rc && fetch) rc = ua_if_fetch(c, checksum);
            break;
        default:
            rc = ua_pr_run(c, d->iterations, device, nullptr);
            if (!rc && fetch) rc = ua_pr_fetch(c, checksum);
            break;
    }
    release(d, c);
    return rc;
}

}  // namespace serve

using namespace serve;

extern "C" {

void* ua_serve_create(void) { return new Server(); }

static int add_dataset(Server* s, Kind k, void* base, int iterations) {
    if (!base) return -1;
    Dataset* d = new Dataset();
    d->kind = k;
    d->base = base;
    d->iterations = iterations;
    s->ds.push_back(d);
    return (int)s->ds.size() - 1;
}

int ua_serve_add_mm(void* h, int n, int block, int precision, uint64_t seed) {
    Server* s = (Server*)h;
    void* mm = ua_mm_create(n, block, precision);
    if (!mm) return -1;
    Rng r(seed);
    std::vector<float> A((size_t)n * n), B((size_t)n * n);
    for (auto& v : A) v = (float)(2.0 * r.uniform() - 1.0);
    for (auto& v : B) v = (float)(2.0 * r.uniform() - 1.0);
    if (ua_mm_set_inputs(mm, A.data(), B.data()) != 0) { ua_mm_destroy(mm); return -1; }
    return add_dataset(s, KIND_MM, mm, 0);
}

int ua_serve_add_if(void* h, int side, int radius, int mode, int shift, int tile, uint64_t seed) {
    Server* s = (Server*)h;
    if (radius < 1 || radius > 15) return -1;
    Rng r(seed);
    ua_if_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.width = cfg.height = side;
    cfg.radius = radius;
    cfg.mode = mode;
    cfg.shift = shift;
    cfg.tile = tile;
    const int d = 2 * radius + 1;
    for (int i = 0; i < d * d; ++i) cfg.weights[i] = (int8_t)((int)(r.next() % 9) - 4);
    void* f = ua_if_create(&cfg);
    if (!f) return -1;
    std::vector<uint8_t> img((size_t)side * side * 3);
    for (size_t i = 0; i < img.size(); i += 8) {
        uint64_t w = r.next();
        for (size_t k = 0; k < 8 && i + k < img.size(); ++k) img[i + k] = (uint8_t)(w >> (8 * k));
    }
    if (ua_if_set_input(f, img.data()) != 0) { ua_if_destroy(f); return -1; }
    return add_dataset(s, KIND_IF, f, 0);
}

int ua_serve_add_pr(void* h, uint32_t pages, uint32_t per_subtask, int iterations, uint64_t seed) {
    Server* s = (Server*)h;
    std::vector<uint32_t> rp((size_t)pages + 1);
    uint64_t edges = 0;
    if (ua_webgen(pages, seed, 10.0, 100, 0.05, rp.data(), nullptr, 0, &edges) != 0) return -1;
    std::vector<uint32_t> col(edges ? edges : 1);
    if (ua_webgen(pages, seed, 10.0, 100, 0.05, rp.data(), col.data(), edges, &edges) != 0) return -1;
    void* p = ua_pr_create(pages, rp.data(), col.data(), per_subtask, 0.85f);
    return add_dataset(s, KIND_PR, p, iterations);
}

int ua_serve_run(void* h, int n, const int* dataset, const double* arrival_s, int slots, int device,
                 int fetch, double* latency_ms, uint64_t* checksums, double* wall_s) {
    Server* s = (Server*)h;
    if (!s || n < 0 || slots < 1 || device < 0 || device > 2) return -1;
    for (int i = 0; i < n; ++i) {
        if (dataset[i] < 0 || dataset[i] >= (int)s->ds.size()) return -1;
        if (i && arrival_s[i] < arrival_s[i - 1]) {
            apps::set_error("ua_serve_run: arrivals must be sorted");
            return -1;
        }
    }
    std::atomic<int> next{0};
    std::atomic<int> failed{0};
    const double t0 = apps::now_secs();
    auto worker = [&]() {
        for (;;) {
            const int i = next++;
            if (i >= n) return;
            const double due = t0 + arrival_s[i];
            double now = apps::now_secs();
            if (now < due) std::this_thread::sleep_for(std::chrono::duration<double>(due - now));
            uint64_t cs = 0;
            const int rc = run_request(s->ds[dataset[i]], device, fetch, checksums ? &cs : nullptr);
            const double done = apps::now_secs();
            if (rc) {
                int expected = 0;
                failed.compare_exchange_strong(expected, rc);
            }
            if (latency_ms) latency_ms[i] = 1e3 * (done - due);
            if (checksums) checksums[i] = cs;
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < slots; ++k) th.emplace_back(worker);
    for (auto& t : th) t.join();
    if (wall_s) *wall_s = apps::now_secs() - t0;
    return failed.load();
}

void* ua_serve_dataset(void* h, int ds) {
    Server* s = (Server*)h;
    if (!s || ds < 0 || ds >= (int)s->ds.size()) return nullptr;
    return s->ds[ds]->base;
}

uint64_t ua_hash(const void* p, uint64_t n) { return apps::hash_bytes(p, (size_t)n); }

int ua_serve_destroy(void* h) {
    Server* s = (Server*)h;
    if (!s) return 0;
    for (Dataset* d : s->ds) {
        for (void* c : d->all_clones) {
            if (d->kind == KIND_MM) ua_mm_destroy(c);
            else if (d->kind == KIND_IF) ua_if_destroy(c);
            else ua_pr_destroy(c);
        }
        if (d->kind == KIND_MM) ua_mm_destroy(d->base);
        else if (d->kind == KIND_IF) ua_if_destroy(d->base);
        else ua_pr_destroy(d->base);
        delete d;
    }
    delete s;
    return 0;
}

}  // extern "C"
