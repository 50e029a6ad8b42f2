// Binned pageRank layout: planning, building and the CPU twin of the device unit.
#include "prbinned.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "unicorn_hip.h"

namespace prb {

static size_t align16(size_t v) { return (v + 15) / 16 * 16; }

static uint32_t unit_count(uint32_t len) { return (len + UH_PRB_UNIT_MAX - 1) / UH_PRB_UNIT_MAX; }

// Page -> unit of a plan.
static std::vector<uint32_t> unit_of_pages(const Layout& L) {
    std::vector<uint32_t> uo(L.pages);
    for (uint32_t u = 0; u < L.units; ++u)
        std::fill(uo.begin() + L.unit_start[u], uo.begin() + L.unit_start[u + 1], u);
    return uo;
}

static bool plan_units(uint32_t N, uint64_t E, uint32_t P, uint32_t cus, Layout& L) {
    L = Layout();
    if (!N || !P) return false;
    L.pages = N;
    L.per_subtask = P;
    L.subtasks = (uint32_t)(((uint64_t)N + P - 1) / P);
    L.links = E;
    const uint32_t S = L.subtasks;
    uint64_t k0 = 0;
    for (uint32_t s = 0; s < S; ++s) k0 += unit_count((uint32_t)std::min<uint64_t>(P, (uint64_t)N - (uint64_t)s * P));
    // Whole rounds of units over the CUs, as long as a run keeps about 128 links or more (a run
    // is the links of one unit into another: there are about units^2 of them).
    uint64_t target = k0;
    if (cus) {
        const uint64_t up = (k0 + cus - 1) / cus * cus;
        const uint64_t cap = (uint64_t)std::sqrt((double)E / 128.0);
        target = std::max(k0, std::min(up, cap));
    }
    L.sub_unit.assign(S + 1, 0);
    L.unit_start.clear();
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t s0 = s * P;
        const uint32_t len = (uint32_t)std::min<uint64_t>(P, (uint64_t)N - s0);
        // cumulative rounding: the shares of the subtasks add up to the target
        const uint64_t share = (uint64_t)std::llround((double)target * ((double)s0 + len) / N) -
                               (uint64_t)std::llround((double)target * (double)s0 / N);
        uint64_t nu = std::max<uint64_t>(unit_count(len), share);
        nu = std::max<uint64_t>(1, std::min<uint64_t>(nu, len));
        L.sub_unit[s] = (uint32_t)L.unit_start.size();
        for (uint64_t j = 0; j < nu; ++j) L.unit_start.push_back(s0 + (uint32_t)((uint64_t)len * j / nu));
    }
    L.sub_unit[S] = (uint32_t)L.unit_start.size();
    L.units = (uint32_t)L.unit_start.size();
    L.unit_start.push_back(N);
    return true;
}

bool plan(uint32_t N, const uint32_t* rp, const uint32_t* col, uint32_t P, uint32_t cus, Layout& L) {
    if (!plan_units(N, rp[N], P, cus, L)) return false;
    const uint32_t K = L.units;
    const std::vector<uint32_t> uo = unit_of_pages(L);
    std::vector<uint64_t> cnt(K, 0);
    std::vector<uint32_t> mark(K, 0xffffffffu);
    uint64_t runs = 0;
    for (uint32_t c = 0; c < K; ++c) {
        for (uint64_t e = rp[L.unit_start[c]]; e < rp[L.unit_start[c + 1]]; ++e) {
            const uint32_t b = uo[col[e]];
            ++cnt[b];
            if (mark[b] != c) { mark[b] = c; ++runs; }
        }
    }
    L.bin_start.assign(K + 1, 0);
    uint64_t slots = 0;
    for (uint32_t b = 0; b < K; ++b) {
        L.bin_start[b] = (uint32_t)slots;
        slots += (cnt[b] + 3) / 4 * 4;
        if (slots > 0xfffffff0ull) return false;
    }
    L.bin_start[K] = (uint32_t)slots;
    if (runs > 0xfffffff0ull) return false;
    L.slots = slots;
    L.runs = (uint32_t)runs;
    size_t o = 0;
    L.off_deg = o;  o = align16(o + (size_t)N * 4);
    L.off_unit = o; o = align16(o + ((size_t)K + 1) * 4);
    L.off_bin = o;  o = align16(o + ((size_t)K + 1) * 4);
    L.off_rptr = o; o = align16(o + ((size_t)K + 1) * 4);
    L.off_roff = o; o = align16(o + ((size_t)L.runs + 1) * 4);
    L.off_rpos = o; o = align16(o + (size_t)L.runs * 4);
    L.off_dst = o;  o = align16(o + (size_t)slots * 2);
    L.off_src = o;  o = align16(o + (size_t)L.links * 2);
    L.graph_bytes = o;
    L.state_values_off = align16((size_t)N * 4);
    L.state_bytes = L.state_values_off + (size_t)slots * 4;
    return true;
}

int build(const uint32_t* rp, const uint32_t* col, Layout& L, uint8_t* graph) {
    const uint32_t N = L.pages, K = L.units, S = L.subtasks;
    if (rp[N] != L.links) return -1;
    memset(graph, 0, L.graph_bytes);
    uint32_t* deg = (uint32_t*)(graph + L.off_deg);
    uint32_t* us = (uint32_t*)(graph + L.off_unit);
    uint32_t* bs = (uint32_t*)(graph + L.off_bin);
    uint32_t* rptr = (uint32_t*)(graph + L.off_rptr);
    uint32_t* roff = (uint32_t*)(graph + L.off_roff);
    uint32_t* rpos = (uint32_t*)(graph + L.off_rpos);
    uint16_t* dst16 = (uint16_t*)(graph + L.off_dst);
    uint16_t* src16 = (uint16_t*)(graph + L.off_src);
    for (uint32_t p = 0; p < N; ++p) deg[p] = rp[p + 1] - rp[p];
    memcpy(us, L.unit_start.data(), ((size_t)K + 1) * 4);
    memcpy(bs, L.bin_start.data(), ((size_t)K + 1) * 4);
    const std::vector<uint32_t> uo = unit_of_pages(L);
    std::vector<uint32_t> next(L.bin_start.begin(), L.bin_start.end() - 1);   // next free slot per bin
    std::vector<uint32_t> cnt(K, 0), cursor(K, 0);
    L.sub_start.assign((size_t)K * (S + 1), 0);
    L.run_ptr.assign(K + 1, 0);
    L.run_off.assign((size_t)L.runs + 1, 0);
    uint32_t r = 0, so = 0;
    uint32_t s = 0;
    for (uint32_t c = 0; c < K; ++c) {
        while (s < S && L.sub_unit[s] == c) {          // first chunk of subtask s: its links start here
            for (uint32_t b = 0; b < K; ++b) L.sub_start[(size_t)b * (S + 1) + s] = next[b];
            ++s;
        }
        const uint64_t e0 = rp[L.unit_start[c]], e1 = rp[L.unit_start[c + 1]];
        for (uint64_t e = e0; e < e1; ++e) ++cnt[uo[col[e]]];
        // the chunk's runs in bin order: sources at src16[so ..], values at the bin's next slots
        L.run_ptr[c] = r;
        for (uint32_t b = 0; b < K; ++b) {
            if (!cnt[b]) continue;
            if (r >= L.runs) return -1;
            L.run_off[r] = so;
            rpos[r] = next[b];
            cursor[b] = so;
            so += cnt[b];
            ++r;
        }
        for (uint32_t src = L.unit_start[c]; src < L.unit_start[c + 1]; ++src) {
            for (uint64_t e = rp[src]; e < rp[src + 1]; ++e) {
                const uint32_t t = col[e];
                const uint32_t b = uo[t];
                src16[cursor[b]++] = (uint16_t)(src - L.unit_start[c]);
                dst16[next[b]++] = (uint16_t)(t - L.unit_start[b]);
            }
        }
        for (uint32_t b = 0; b < K; ++b) cnt[b] = 0;
    }
    L.run_ptr[K] = r;
    L.run_off[r] = so;
    if (r != L.runs || so != L.links) return -1;
    for (uint32_t b = 0; b < K; ++b) {
        L.sub_start[(size_t)b * (S + 1) + S] = next[b];
        if (next[b] > L.bin_start[b + 1] || L.bin_start[b + 1] - next[b] > 3) return -1;
    }
    memcpy(rptr, L.run_ptr.data(), ((size_t)K + 1) * 4);
    memcpy(roff, L.run_off.data(), ((size_t)L.runs + 1) * 4);
    return 0;
}

uint64_t to_fixed(float v, int s) {
    uint32_t b;
    memcpy(&b, &v, 4);
    int e = (int)(b >> 23);
    uint64_t m = b & 0x7fffffu;
    if (e) m |= 0x800000u; else e = 1;
    const int sh = e - 150 + s;
    return sh >= 0 ? (m << sh) : (sh > -64 ? (m >> -sh) : 0ull);
}

int scale_exp(uint32_t pages, float x0) {
    const double bound = (double)pages * (double)x0 + 1.0;
    int bits = (int)std::ceil(std::log2(bound));
    if (bits < 1) bits = 1;
    return 62 - bits;
}

void unit_cpu(const Views& v, uint32_t u, int kind, float damping, float x0, int s, std::vector<uint64_t>& acc,
              std::vector<float>& w) {
    const uint32_t p0 = v.unit_start[u], rows = v.unit_start[u + 1] - p0;
    w.assign(rows, 0.f);
    if (kind == 0) {
        for (uint32_t r = 0; r < rows; ++r) {
            v.y[p0 + r] = x0;
            const uint32_t d = v.deg[p0 + r];
            w[r] = d ? (damping * x0) / (float)d : 0.f;
        }
    } else {
        acc.assign(rows, 0);
        for (uint32_t k = v.bin_start[u]; k < v.bin_start[u + 1]; ++k) acc[v.dst16[k]] += to_fixed(v.values_in[k], s);
        const double unscale = std::ldexp(1.0, -s);
        for (uint32_t r = 0; r < rows; ++r) {
            const float y = (float)((double)(long long)acc[r] * unscale);
            v.y[p0 + r] = y;
            const uint32_t d = v.deg[p0 + r];
            w[r] = d ? (damping * y) / (float)d : 0.f;
        }
    }
    if (!v.values_out) return;
    for (uint32_t r = v.run_ptr[u]; r < v.run_ptr[u + 1]; ++r) {
        const uint32_t o0 = v.run_off[r], o1 = v.run_off[r + 1], pos = v.run_pos[r];
        for (uint32_t k = o0; k < o1; ++k) v.values_out[pos + (k - o0)] = w[v.src16[k]];
    }
}

}  // namespace prb
