// Synthetic power-law web graph for pageRank (the role of the reference's
// testSuite/pageRank/webGenerator.pl, with the distribution the benchmark configuration names).
//
//   out-degree: 0 with probability `dangling`; otherwise k in [1, max_out] with
//               P(k) proportional to (k + k0)^-2, k0 solved so that the overall mean is
//               `mean_out` (a plain Zipf(2) truncated to [1,100] has mean ~3.2, so the shift is
//               what lets "s = 2, mean 10" hold);
//   targets:    a Polya urn over pages that starts with one ticket per page and gains one
//               ticket per drawn target, i.e. preferential attachment by in-degree + 1;
//               no self links and no duplicate targets within a page (redrawn).
// Degrees and targets come from two independent xoshiro256** streams of one seed, so the
// degree pass (which sizes the edge array) and the target pass are reproducible separately.
#include "app_common.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace webgen {

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {   // splitmix64
            z += 0x9e3779b97f4a7c15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
};

static double shifted_zipf_mean(double k0, int max_out) {
    double num = 0, den = 0;
    for (int k = 1; k <= max_out; ++k) {
        const double p = 1.0 / ((k + k0) * (k + k0));
        num += k * p;
        den += p;
    }
    return num / den;
}

double solve_shift(double conditional_mean, int max_out) {
    double lo = -0.999, hi = 1e6;
    if (shifted_zipf_mean(lo, max_out) >= conditional_mean) return lo;
    if (shifted_zipf_mean(hi, max_out) <= conditional_mean) return hi;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (shifted_zipf_mean(mid, max_out) < conditional_mean) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace webgen

extern "C" double ua_webgen_shift(double mean_out, uint32_t max_out, double dangling) {
    return webgen::solve_shift(mean_out / (1.0 - dangling), (int)max_out);
}

// Phase 1 (col == NULL): fill row_ptr[0..pages] and *edges. Phase 2: also fill col (0-based
// target ids, capacity col_capacity).
extern "C" int ua_webgen(uint32_t pages, uint64_t seed, double mean_out, uint32_t max_out, double dangling,
                         uint32_t* row_ptr, uint32_t* col, uint64_t col_capacity, uint64_t* edges) {
    if (!pages || !row_ptr || max_out == 0 || dangling < 0 || dangling >= 1 || mean_out <= 0) {
        apps::set_error("ua_webgen: invalid arguments");
        return -1;
    }
    const double k0 = webgen::solve_shift(mean_out / (1.0 - dangling), (int)max_out);
    std::vector<double> cdf(max_out);
    double acc = 0;
    for (uint32_t k = 1; k <= max_out; ++k) {
        acc += 1.0 / ((k + k0) * (k + k0));
        cdf[k - 1] = acc;
    }
    for (auto& c : cdf) c /= acc;
    webgen::Xoshiro rd(seed * 2 + 1);
    uint64_t e = 0;
    row_ptr[0] = 0;
    for (uint32_t p = 0; p < pages; ++p) {
        uint32_t k = 0;
        if (rd.uniform() >= dangling) {
            const double u = rd.uniform();
            k = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
            if (k > max_out) k = max_out;
        }
        if (pages > 1 && k > pages - 1) k = pages - 1;
        if (pages == 1) k = 0;
        e += k;
        if (e > 0xffffffffull) { apps::set_error("ua_webgen: more than 2^32 edges"); return -1; }
        row_ptr[p + 1] = (uint32_t)e;
    }
    *edges = e;
    if (!col) return 0;
    if (col_capacity < e) { apps::set_error("ua_webgen: col capacity"); return -1; }
    std::vector<uint32_t> urn;
    urn.reserve((size_t)pages + e);
    for (uint32_t p = 0; p < pages; ++p) urn.push_back(p);
    webgen::Xoshiro rt(seed * 2 + 2);
    for (uint32_t p = 0; p < pages; ++p) {
        const uint32_t b = row_ptr[p], n = row_ptr[p + 1] - b;
        for (uint32_t j = 0; j < n; ++j) {
            auto taken = [&](uint32_t cand) {
                if (cand == p) return true;
                for (uint32_t q = 0; q < j; ++q) if (col[b + q] == cand) return true;
                return false;
            };
            uint32_t t = 0;
            bool ok = false;
            for (int tries = 0; tries < 64 && !ok; ++tries) {
                t = urn[rt.below(urn.size())];
                ok = !taken(t);
            }
            if (!ok) {   // fall back to the next free page id (one exists: k <= pages - 1)
                for (uint32_t c = 0; c < pages; ++c) {
                    const uint32_t cand = (t + 1 + c) % pages;
                    if (!taken(cand)) { t = cand; break; }
                }
            }
            col[b + j] = t;
            urn.push_back(t);
        }
    }
    return 0;
}
