// pageRank app on the pm:: API (reference testSuite/pageRank/source/code/src/pageRank.cpp and
// pageRankCUDA.cu).
//
// Recurrence (reference pageRank.cpp:237-276, 336-378): on iteration 0 every page has rank 1.0;
// each iteration, y[t] = sum over links src -> t of (0.85 * x[src]) / outdeg(src); no teleport
// term, dangling pages contribute nothing.
//
// Decomposition. The reference splits the *source* pages into subtasks, each pushing into a
// private full-length output vector that a float REDUCE_ADD then folds on the host, and streams
// the padded web dump from mmap'd files into the GPU every iteration. Here the graph is an
// address space holding an in-link CSR (built once at load), subtasks own *destination* page
// ranges, and each computes its rows by pulling from the sources' increments: no atomics, no
// partial vectors, no reduction, and the graph stays resident in HBM across iterations.
// Address spaces of one iteration task:
//   0 graph  (READ_ONLY):  [in_row_ptr (N+1) | outdeg (N) | in_col (E) | bseg | bcnt]
//            uint32 except bcnt (uint8). When every page has at most 255 in-links from each of
//            the UH_PR_BLOCKS source blocks, in_col is kept in the source-blocked order of
//            uh_pagerank_args (per row group of UH_PR_GROUP_ROWS pages: by source block, then
//            destination, then source) with its segment starts bseg and per-row counts bcnt;
//            otherwise in_col is the plain in-link CSR and bseg / bcnt are absent.
//   1 state  (READ_ONLY):  [x (N) | w (N)] float, w[p] = (0.85 * x[p]) / outdeg(p)
//   2 state' (WRITE_ONLY): [y (N) | w' (N)] float, the next state
// Subtask s subscribes to rows [s0, s1) of the graph (row pointers, degrees and their in-link
// range; in the blocked layout the in-links, bseg and bcnt of the row groups that overlap
// [s0, s1)), to all of w, and writes y[s0:s1) and w'[s0:s1). An "init" task writes the
// iteration-0 state (x = 1, w = 0.85 / outdeg).
#include "app_common.h"
#include "prbinned.h"
#include "unicorn_hip.h"
#include "unicorn_apps.h"

#include <algorithm>
#include <vector>

namespace pageRank {

using namespace pm;

enum memIndex { GRAPH_MEM_INDEX = 0, STATE_IN_MEM_INDEX = 1, STATE_OUT_MEM_INDEX = 2 };

enum Layout { LAYOUT_CSR = 0, LAYOUT_BLOCKED = 1, LAYOUT_BINNED = 2 };

struct pageRankTaskConf {
    uint32_t totalWebPages;
    uint32_t webPagesPerSubtask;
    uint64_t totalLinks;
    float dampingFactor;
    float initialPageRank;
    int kind;                        // 0 = init, 1 = iterate
    const uint32_t* hostRowPtr;      // in-link row pointers on the host (subscription sizing)
    int blocked;                     // graph in the source-blocked layout
    const uint32_t* hostBseg;        // its segment starts on the host (subscription sizing)
    int layout;                      // Layout
    int produceNext;                 // binned: write the next iteration's link values
    int scaleExp;                    // binned: fixed-point exponent of the row sums
    const prb::Layout* hostBinned;   // binned: host tables (subscription sizing)
};

static constexpr uint32_t GR = UH_PR_GROUP_ROWS;
static constexpr uint32_t NB = UH_PR_BLOCKS;

static inline uint32_t groups_of(uint64_t pages) { return (uint32_t)((pages + GR - 1) / GR); }
// Pages per source block: ceil(N / NB) rounded up to 32 pages (a 128-byte line of w).
static inline uint32_t block_pages(uint64_t pages) { return (uint32_t)(((pages + NB - 1) / NB + 31) / 32 * 32); }
static inline size_t bcnt_index(uint32_t t, uint32_t x) { return ((size_t)(t / GR) * NB + x) * GR + t % GR; }

static inline size_t row_ptr_off(uint32_t t) { return (size_t)t * 4; }
static inline size_t deg_off(const pageRankTaskConf* c, uint32_t t) { return ((size_t)c->totalWebPages + 1 + t) * 4; }
static inline size_t col_off(const pageRankTaskConf* c, uint64_t e) { return ((size_t)2 * c->totalWebPages + 1 + e) * 4; }
static inline size_t bseg_off(const pageRankTaskConf* c, size_t i) {
    return ((size_t)2 * c->totalWebPages + 1 + c->totalLinks + i) * 4;
}
static inline size_t bcnt_off(const pageRankTaskConf* c, size_t j) {
    return bseg_off(c, (size_t)groups_of(c->totalWebPages) * NB + 1) + j;
}
// Row groups overlapping [s0, s1) and their link range in the blocked layout.
static inline void group_range(const pageRankTaskConf* c, uint32_t s0, uint32_t s1, uint32_t* g0, uint32_t* ng,
                               uint64_t* e0, uint64_t* e1) {
    *g0 = s0 / GR;
    *ng = (s1 - 1) / GR - *g0 + 1;
    *e0 = c->hostBseg[(size_t)*g0 * NB];
    *e1 = c->hostBseg[(size_t)(*g0 + *ng) * NB];
}

static void range_of(const pageRankTaskConf* c, ulong id, uint32_t* s0, uint32_t* s1) {
    *s0 = (uint32_t)(id * c->webPagesPerSubtask);
    const uint64_t e = (uint64_t)*s0 + c->webPagesPerSubtask;
    *s1 = (uint32_t)(e < c->totalWebPages ? e : c->totalWebPages);
}

// ---- binned layout (prbinned.h): the records a subtask subscribes, and pointers into them ----

static inline size_t down16(size_t v) { return v & ~(size_t)15; }
static inline size_t up16(size_t v) { return (v + 15) & ~(size_t)15; }

struct Rec { size_t off, len; };

// Read records of the graph for subtask s (each rounded out to 16 bytes, so that a COMPACT view
// keeps every region's 16-byte alignment), in address order.
static std::vector<Rec> binned_graph_records(const pageRankTaskConf* c, uint32_t s) {
    const prb::Layout& L = *c->hostBinned;
    uint32_t s0, s1;
    range_of(c, s, &s0, &s1);
    const uint32_t u0 = L.sub_unit[s], u1 = L.sub_unit[s + 1];
    std::vector<Rec> r;
    auto add = [&](size_t off, size_t len) {
        if (len) r.push_back({down16(off), up16(off + len) - down16(off)});
    };
    add(L.off_deg + (size_t)s0 * 4, (size_t)(s1 - s0) * 4);
    add(L.off_unit + (size_t)u0 * 4, ((size_t)u1 - u0 + 1) * 4);
    if (c->kind == 1) add(L.off_bin + (size_t)u0 * 4, ((size_t)u1 - u0 + 1) * 4);
    if (c->produceNext) {
        const uint32_t r0 = L.run_ptr[u0], r1 = L.run_ptr[u1];
        add(L.off_rptr + (size_t)u0 * 4, ((size_t)u1 - u0 + 1) * 4);
        add(L.off_roff + (size_t)r0 * 4, ((size_t)r1 - r0 + 1) * 4);
        add(L.off_rpos + (size_t)r0 * 4, ((size_t)r1 - r0) * 4);
    }
    if (c->kind == 1) add(L.off_dst + (size_t)L.bin_start[u0] * 2, ((size_t)L.bin_start[u1] - L.bin_start[u0]) * 2);
    if (c->produceNext) {
        const uint32_t o0 = L.run_off[L.run_ptr[u0]], o1 = L.run_off[L.run_ptr[u1]];
        add(L.off_src + (size_t)o0 * 2, ((size_t)o1 - o0) * 2);
    }
    return r;
}

// Pointer to address-space offset `off` inside the view of records `recs` (address order).
static uint8_t* rec_ptr(const pmMemInfo& mi, const std::vector<Rec>& recs, size_t off) {
    uint8_t* base = (uint8_t*)mi.ptr;
    if (mi.visibilityType == SUBSCRIPTION_NATURAL) return base + (off - recs.front().off);
    size_t packed = 0;
    for (const Rec& r : recs) {
        if (off >= r.off && off < r.off + r.len) return base + packed + (off - r.off);
        packed += r.len;
    }
    return nullptr;
}

static void binned_subscribe(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si, const pageRankTaskConf* c) {
    const prb::Layout& L = *c->hostBinned;
    const uint32_t s = (uint32_t)si.subtaskId, S = L.subtasks;
    uint32_t s0, s1;
    range_of(c, s, &s0, &s1);
    for (const Rec& r : binned_graph_records(c, s))
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX,
                            READ_SUBSCRIPTION, pmSubscriptionInfo(r.off, r.len));
    const uint32_t u0 = L.sub_unit[s], u1 = L.sub_unit[s + 1];
    if (c->kind == 1 && L.bin_start[u1] > L.bin_start[u0])
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, STATE_IN_MEM_INDEX,
                            READ_SUBSCRIPTION,
                            pmSubscriptionInfo(L.state_values_off + (size_t)L.bin_start[u0] * 4,
                                               ((size_t)L.bin_start[u1] - L.bin_start[u0]) * 4));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, STATE_OUT_MEM_INDEX,
                        WRITE_SUBSCRIPTION, pmSubscriptionInfo((size_t)s0 * 4, (size_t)(s1 - s0) * 4));
    if (c->produceNext) {
        // the values of this subtask's chunks: one contiguous range in every bin
        for (uint32_t b = 0; b < L.units; ++b) {
            const size_t v0 = L.sub_start[(size_t)b * (S + 1) + s], v1 = L.sub_start[(size_t)b * (S + 1) + s + 1];
            if (v1 > v0)
                pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, STATE_OUT_MEM_INDEX,
                                    WRITE_SUBSCRIPTION, pmSubscriptionInfo(L.state_values_off + v0 * 4, (v1 - v0) * 4));
        }
    }
}

// Global-indexed pointers of a binned subtask (device or host memory alike).
static prb::Views binned_views(const pageRankTaskConf* c, const pmSubtaskInfo& si) {
    const prb::Layout& L = *c->hostBinned;
    const uint32_t s = (uint32_t)si.subtaskId;
    const uint32_t u0 = L.sub_unit[s];
    uint32_t s0, s1;
    range_of(c, s, &s0, &s1);
    const std::vector<Rec> recs = binned_graph_records(c, s);
    const pmMemInfo& g = si.memInfo[GRAPH_MEM_INDEX];
    prb::Views v;
    memset(&v, 0, sizeof(v));
    v.deg = (const uint32_t*)rec_ptr(g, recs, L.off_deg + (size_t)s0 * 4) - s0;
    v.unit_start = (const uint32_t*)rec_ptr(g, recs, L.off_unit + (size_t)u0 * 4) - u0;
    if (c->kind == 1) {
        v.bin_start = (const uint32_t*)rec_ptr(g, recs, L.off_bin + (size_t)u0 * 4) - u0;
        const uint32_t k0 = L.bin_start[u0];
        if (L.bin_start[L.sub_unit[s + 1]] > k0) {
            v.dst16 = (const uint16_t*)rec_ptr(g, recs, L.off_dst + (size_t)k0 * 2) - k0;
            v.values_in = (const float*)si.memInfo[STATE_IN_MEM_INDEX].ptr - k0;   // one record
        }
    }
    if (c->produceNext) {
        const uint32_t r0 = L.run_ptr[u0];
        v.run_ptr = (const uint32_t*)rec_ptr(g, recs, L.off_rptr + (size_t)u0 * 4) - u0;
        v.run_off = (const uint32_t*)rec_ptr(g, recs, L.off_roff + (size_t)r0 * 4) - r0;
        if (L.run_ptr[L.sub_unit[s + 1]] > r0) {
            v.run_pos = (const uint32_t*)rec_ptr(g, recs, L.off_rpos + (size_t)r0 * 4) - r0;
            const uint32_t o0 = L.run_off[r0];
            v.src16 = (const uint16_t*)rec_ptr(g, recs, L.off_src + (size_t)o0 * 2) - o0;
        }
    }
    // the output state is subscribed with a NATURAL view: ptr is the offset of y[s0]
    uint8_t* ob = (uint8_t*)si.memInfo[STATE_OUT_MEM_INDEX].ptr - (size_t)s0 * 4;
    v.y = (float*)ob;
    v.values_out = c->produceNext ? (float*)(ob + L.state_values_off) : nullptr;
    return v;
}

static pmStatus prDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    if (c->layout == LAYOUT_BINNED) {
        binned_subscribe(ti, di, si, c);
        return pmSuccess;
    }
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    const size_t N = c->totalWebPages;
    // graph rows: row pointers s0..s1 (inclusive), degrees s0..s1-1 (the init task needs only these)
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX, READ_SUBSCRIPTION,
                        pmSubscriptionInfo(row_ptr_off(s0), (size_t)(s1 - s0 + 1) * 4));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX, READ_SUBSCRIPTION,
                        pmSubscriptionInfo(deg_off(c, s0), (size_t)(s1 - s0) * 4));
    if (c->kind == 1 && c->blocked) {
        uint32_t g0, ng;
        uint64_t e0, e1;
        group_range(c, s0, s1, &g0, &ng, &e0, &e1);
        if (e1 > e0)
            pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX,
                                READ_SUBSCRIPTION, pmSubscriptionInfo(col_off(c, e0), (size_t)(e1 - e0) * 4));
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX, READ_SUBSCRIPTION,
                            pmSubscriptionInfo(bseg_off(c, (size_t)g0 * NB), ((size_t)ng * NB + 1) * 4));
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX, READ_SUBSCRIPTION,
                            pmSubscriptionInfo(bcnt_off(c, (size_t)g0 * NB * GR), (size_t)ng * NB * GR));
    } else if (c->kind == 1) {
        const uint64_t e0 = c->hostRowPtr[s0], e1 = c->hostRowPtr[s1];
        if (e1 > e0)
            pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, GRAPH_MEM_INDEX,
                                READ_SUBSCRIPTION, pmSubscriptionInfo(col_off(c, e0), (size_t)(e1 - e0) * 4));
    }
    if (c->kind == 1) {
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, STATE_IN_MEM_INDEX,
                            READ_SUBSCRIPTION, pmSubscriptionInfo(N * 4, N * 4));
    }
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, STATE_OUT_MEM_INDEX, WRITE_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo((size_t)s0 * 4, (size_t)(s1 - s0) * 4, N * 4, 2));
    return pmSuccess;
}

// Device/host pointers of one subtask, rebased so that they are indexed by global page id
// (and edge id for col), for NATURAL and COMPACT views alike.
struct Views {
    const uint32_t* row_ptr;
    const uint32_t* deg;
    const uint32_t* col;
    const uint32_t* bseg;     // blocked layout: indexed by g * NB + x
    const uint8_t* bcnt;      // blocked layout: indexed by (g * NB + x) * GR + r
    const float* w;
    float* y;
    float* w_next;
};

static Views views_of(const pageRankTaskConf* c, const pmSubtaskInfo& si, uint32_t s0, uint32_t s1) {
    Views v;
    const pmMemInfo& g = si.memInfo[GRAPH_MEM_INDEX];
    const pmMemInfo& in = si.memInfo[STATE_IN_MEM_INDEX];
    const pmMemInfo& out = si.memInfo[STATE_OUT_MEM_INDEX];
    const size_t N = c->totalWebPages;
    const bool blk = c->kind == 1 && c->blocked;
    uint32_t g0 = 0, ng = 0;
    uint64_t e0 = c->hostRowPtr[s0], e1 = c->hostRowPtr[s1];
    if (blk) group_range(c, s0, s1, &g0, &ng, &e0, &e1);
    const size_t b0 = (size_t)g0 * NB;          // first bseg entry of the subtask's groups
    const uint8_t* gb = (const uint8_t*)g.ptr;
    v.bseg = nullptr;
    v.bcnt = nullptr;
    if (g.visibilityType == SUBSCRIPTION_NATURAL) {
        // ptr is the start of the first subscribed record (row_ptr[s0]); distances are preserved
        const size_t o = row_ptr_off(s0);
        v.row_ptr = (const uint32_t*)gb - s0;
        v.deg = (const uint32_t*)(gb + (deg_off(c, s0) - o)) - s0;
        v.col = (const uint32_t*)(gb + (col_off(c, e0) - o)) - e0;
        if (blk) {
            v.bseg = (const uint32_t*)(gb + (bseg_off(c, b0) - o)) - b0;
            v.bcnt = gb + (bcnt_off(c, b0 * GR) - o) - b0 * GR;
        }
    } else {
        // packed records in address order: row pointers, degrees, in-links (, bseg, bcnt)
        const size_t rp = (size_t)(s1 - s0 + 1) * 4, dg = (size_t)(s1 - s0) * 4;
        v.row_ptr = (const uint32_t*)gb - s0;
        v.deg = (const uint32_t*)(gb + rp) - s0;
        v.col = (const uint32_t*)(gb + rp + dg) - e0;
        if (blk) {
            const size_t cb = (size_t)(e1 - e0) * 4, sb = ((size_t)ng * NB + 1) * 4;
            v.bseg = (const uint32_t*)(gb + rp + dg + cb) - b0;
            v.bcnt = gb + rp + dg + cb + sb - b0 * GR;
        }
    }
    v.w = nullptr;
    if (c->kind == 1) {
        // state_in: one record, the whole w vector
        v.w = (const float*)in.ptr;
    }
    uint8_t* ob = (uint8_t*)out.ptr;
    if (out.visibilityType == SUBSCRIPTION_NATURAL) {
        v.y = (float*)ob - s0;
        v.w_next = (float*)(ob + N * 4) - s0;
    } else {
        v.y = (float*)ob - s0;
        v.w_next = (float*)(ob + (size_t)(s1 - s0) * 4) - s0;
    }
    return v;
}

static pmStatus pr_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    if (c->layout == LAYOUT_BINNED) {
        // the device units' arithmetic, unit by unit (the same bits as the GPU)
        const prb::Layout& L = *c->hostBinned;
        const prb::Views v = binned_views(c, si);
        std::vector<uint64_t> acc;
        std::vector<float> w;
        for (uint32_t u = L.sub_unit[si.subtaskId]; u < L.sub_unit[si.subtaskId + 1]; ++u)
            prb::unit_cpu(v, u, c->kind, c->dampingFactor, c->initialPageRank, c->scaleExp, acc, w);
        return pmSuccess;
    }
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    Views v = views_of(c, si, s0, s1);
    if (c->kind == 1 && c->blocked) {
        // the GPU kernels' order: per row, each block's links in link order, blocks in order
        std::vector<float> y(s1 - s0, 0.f);
        for (uint32_t g = s0 / GR; g <= (s1 - 1) / GR; ++g) {
            for (uint32_t x = 0; x < NB; ++x) {
                const uint8_t* cnt = v.bcnt + ((size_t)g * NB + x) * GR;
                uint32_t k = v.bseg[(size_t)g * NB + x];
                for (uint32_t r = 0; r < GR; ++r) {
                    const uint32_t t = g * GR + r;
                    float p = 0.f;
                    for (uint32_t q = 0; q < cnt[r]; ++q) p += v.w[v.col[k + q]];
                    k += cnt[r];
                    if (t >= s0 && t < s1) y[t - s0] += p;
                }
            }
        }
        for (uint32_t t = s0; t < s1; ++t) {
            v.y[t] = y[t - s0];
            const uint32_t d = v.deg[t];
            v.w_next[t] = d ? (c->dampingFactor * y[t - s0]) / (float)d : 0.f;
        }
        return pmSuccess;
    }
    for (uint32_t t = s0; t < s1; ++t) {
        float y;
        if (c->kind == 0) {
            y = c->initialPageRank;
        } else {
            y = 0.f;
            for (uint32_t e = v.row_ptr[t]; e < v.row_ptr[t + 1]; ++e) y += v.w[v.col[e]];
        }
        v.y[t] = y;
        const uint32_t d = v.deg[t];
        v.w_next[t] = d ? (c->dampingFactor * y) / (float)d : 0.f;
    }
    return pmSuccess;
}

static pmStatus pr_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    if (c->layout == LAYOUT_BINNED) {
        const prb::Layout& L = *c->hostBinned;
        const prb::Views v = binned_views(c, si);
        uh_pagerank_binned_args a;
        a.unit_start = v.unit_start;
        a.bin_start = v.bin_start;
        a.run_ptr = v.run_ptr;
        a.run_off = v.run_off;
        a.run_pos = v.run_pos;
        a.dst16 = v.dst16;
        a.src16 = v.src16;
        a.deg = v.deg;
        a.values_in = v.values_in;
        a.values_out = v.values_out;
        a.y = v.y;
        a.unit0 = L.sub_unit[si.subtaskId];
        a.units = L.sub_unit[si.subtaskId + 1] - a.unit0;
        a.damping = c->dampingFactor;
        a.x0 = c->initialPageRank;
        a.kind = c->kind;
        a.scale_exp = c->scaleExp;
        if (uh_pagerank_binned(&a, stream)) {
            apps::set_error(std::string("pageRank GPU callback: ") + uh_last_error());
            return pmGraphicsCardError;
        }
        return pmSuccess;
    }
    uint32_t s0, s1;
    range_of(c, si.subtaskId, &s0, &s1);
    Views v = views_of(c, si, s0, s1);
    int rc;
    if (c->kind == 0) {
        // the init task is tiny and runs once per chain: y = x0 on the device, w = damping x0 / deg
        rc = uh_fill_u32(v.y + s0, *(const uint32_t*)&c->initialPageRank, s1 - s0, stream);
        if (!rc) rc = uh_pagerank_init_w(v.w_next + s0, v.deg + s0, s1 - s0, c->initialPageRank, c->dampingFactor, stream);
    } else {
        uh_pagerank_args a;
        a.row_ptr = v.row_ptr;
        a.col = v.col;
        a.deg = v.deg;
        a.w = v.w;
        a.y = v.y;
        a.w_next = v.w_next;
        a.row0 = s0;
        a.rows = s1 - s0;
        a.damping = c->dampingFactor;
        a.pad = 0;
        a.bcnt = v.bcnt;
        a.bseg = v.bseg;
        rc = uh_pagerank_pull(&a, stream);
    }
    if (rc) {
        apps::set_error(std::string("pageRank GPU callback: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    uint32_t pages = 0, per_subtask = 0;
    uint64_t edges = 0;
    float damping = 0.85f;
    int layout = LAYOUT_CSR;
    std::shared_ptr<std::vector<uint32_t>> in_row_ptr;   // host copy for subscription sizing (CSR, blocked)
    std::shared_ptr<std::vector<uint32_t>> bseg;         // host copy (blocked layout), or null
    std::shared_ptr<prb::Layout> binned;                 // host tables (binned layout), or null
    size_t state_bytes = 0;
    std::shared_ptr<apps::SharedMem> inputs;             // the graph, shared with clones
    pmMemHandle graph = nullptr, state[2] = {nullptr, nullptr};
    int current = 0;                    // state holding the latest ranks
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;
static int g_layout = LAYOUT_BINNED;    // layout of graphs created from now on (falling back when it cannot hold one)

static const float kInitialPageRank = 1.0f;

static int run_kind(Instance* in, int kind, int device, double* secs, bool produce_next) {
    pageRankTaskConf conf;
    conf.totalWebPages = in->pages;
    conf.webPagesPerSubtask = in->per_subtask;
    conf.totalLinks = in->edges;
    conf.dampingFactor = in->damping;
    conf.initialPageRank = kInitialPageRank;
    conf.kind = kind;
    conf.hostRowPtr = in->in_row_ptr ? in->in_row_ptr->data() : nullptr;
    conf.blocked = in->layout == LAYOUT_BLOCKED ? 1 : 0;
    conf.hostBseg = in->bseg ? in->bseg->data() : nullptr;
    conf.layout = in->layout;
    conf.produceNext = produce_next ? 1 : 0;
    conf.scaleExp = prb::scale_exp(in->pages, kInitialPageRank);
    conf.hostBinned = in->binned.get();
    const int src = in->current, dst = kind == 0 ? in->current : 1 - in->current;
    pmTaskMem mem[3];
    mem[GRAPH_MEM_INDEX] = pmTaskMem(in->graph, READ_ONLY, SUBSCRIPTION_OPTIMAL);
    mem[STATE_IN_MEM_INDEX] = pmTaskMem(in->state[kind == 0 ? 1 - dst : src], READ_ONLY, SUBSCRIPTION_OPTIMAL);
    // binned: a subtask writes its y and one range of link values in every bin; kernels index them
    // by page and by link slot, which only a NATURAL view keeps linear
    mem[STATE_OUT_MEM_INDEX] = pmTaskMem(in->state[dst], WRITE_ONLY,
                                         in->layout == LAYOUT_BINNED ? SUBSCRIPTION_NATURAL : SUBSCRIPTION_OPTIMAL);
    const ulong subtasks = (in->pages + in->per_subtask - 1) / in->per_subtask;
    pmTaskDetails d(&conf, sizeof(conf), mem, 3, g_cb.h[device], subtasks);
    int rc = apps::run_task(d, secs);
    if (rc == 0) in->current = dst;
    return rc;
}

}  // namespace pageRank

using namespace pageRank;

// bcnt of the source-blocked layout (groups * NB * GR bytes) from the out-link CSR. Returns 1, or
// 0 if some page has more than 255 in-links from one source block (the layout cannot hold it).
extern "C" int ua_pr_block_counts(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                                  uint8_t* bcnt) {
    // the blocked kernels address w with 32-bit byte offsets: graphs of 2^30 pages or more keep
    // the plain CSR (checked before any array is read)
    if (pages >= (1u << 30)) return 0;
    const uint32_t bs = block_pages(pages);
    memset(bcnt, 0, (size_t)groups_of(pages) * NB * GR);
    for (uint32_t s = 0; s < pages; ++s) {
        const uint32_t x = s / bs;
        for (uint32_t e = out_row_ptr[s]; e < out_row_ptr[s + 1]; ++e) {
            uint8_t& n = bcnt[bcnt_index(out_col[e], x)];
            if (n == 255) return 0;
            ++n;
        }
    }
    return 1;
}

// Bytes of the graph address space.
extern "C" uint64_t ua_pr_graph_bytes(uint32_t pages, uint64_t edges, int blocked) {
    uint64_t b = (2ull * pages + 1 + edges) * 4;
    if (blocked) b += ((uint64_t)groups_of(pages) * NB + 1) * 4 + (uint64_t)groups_of(pages) * NB * GR;
    return b;
}

// Fills the graph address space from the out-link CSR: the in-link CSR by a counting sort over
// targets (sources ascending within each row) and, if bcnt is given (from ua_pr_block_counts),
// its source-blocked reordering with bseg and bcnt.
extern "C" int ua_pr_build_graph(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                                 const uint8_t* bcnt, uint32_t* graph) {
    const size_t N = pages;
    const uint64_t E = out_row_ptr[pages];
    uint32_t* rp = graph;
    uint32_t* deg = graph + N + 1;
    uint32_t* col = graph + 2 * N + 1;
    std::vector<uint32_t> cnt(N + 1, 0);
    for (uint64_t e = 0; e < E; ++e) cnt[out_col[e] + 1]++;
    rp[0] = 0;
    for (size_t t = 0; t < N; ++t) rp[t + 1] = rp[t] + cnt[t + 1];
    std::vector<uint32_t> pos(rp, rp + N);
    for (uint32_t s = 0; s < pages; ++s) {
        deg[s] = out_row_ptr[s + 1] - out_row_ptr[s];
        for (uint32_t e = out_row_ptr[s]; e < out_row_ptr[s + 1]; ++e) col[pos[out_col[e]]++] = s;
    }
    if (!bcnt) return 0;
    const uint32_t G = groups_of(pages);
    uint32_t* bseg = col + E;
    memcpy((uint8_t*)(bseg + (size_t)G * NB + 1), bcnt, (size_t)G * NB * GR);
    // within each row group: block by block, each row's block-x links (a run of its sorted
    // sources) in row order
    std::vector<uint32_t> tmp, cur(GR);
    for (uint32_t g = 0; g < G; ++g) {
        const uint32_t r0 = g * GR, r1 = (uint32_t)std::min<size_t>(N, (size_t)r0 + GR);
        const uint32_t l0 = rp[r0], l1 = rp[r1];
        tmp.assign(col + l0, col + l1);
        for (uint32_t r = r0; r < r1; ++r) cur[r - r0] = rp[r] - l0;
        uint32_t k = l0;
        for (uint32_t x = 0; x < NB; ++x) {
            bseg[(size_t)g * NB + x] = k;
            const uint8_t* c = bcnt + ((size_t)g * NB + x) * GR;
            for (uint32_t r = r0; r < r1; ++r) {
                const uint32_t n = c[r - r0];
                memcpy(col + k, tmp.data() + cur[r - r0], (size_t)n * 4);
                cur[r - r0] += n;
                k += n;
            }
        }
        if (k != l1) return -1;
    }
    bseg[(size_t)G * NB] = (uint32_t)E;
    return 0;
}

// Layout of the graphs that ua_pr_create builds from now on: 1 = source-blocked where possible
// (the default), 0 = plain in-link CSR.
extern "C" void ua_pr_set_layout(int layout) {
    g_layout = layout == LAYOUT_CSR ? LAYOUT_CSR : layout == LAYOUT_BLOCKED ? LAYOUT_BLOCKED : LAYOUT_BINNED;
}
extern "C" int ua_pr_blocked(void* h) { return ((Instance*)h)->layout == LAYOUT_BLOCKED ? 1 : 0; }
extern "C" int ua_pr_layout(void* h) { return ((Instance*)h)->layout; }

// The binned layout outside a task (for inspection): with graph == NULL, plans and returns the
// graph address space's bytes and fills info = {units, runs, slots, subtasks, region offsets
// (deg, unit_start, bin_start, run_ptr, run_off, run_pos, dst16, src16)}; with a graph buffer of
// that size, also fills it. Returns 0 when the layout cannot hold the graph.
extern "C" uint64_t ua_prb_layout(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                                  uint32_t pages_per_subtask, uint32_t cus, void* graph, uint64_t* info) {
    prb::Layout L;
    if (!pages || !out_row_ptr || !pages_per_subtask ||
        !prb::plan(pages, out_row_ptr, out_col, pages_per_subtask, cus, L))
        return 0;
    if (info) {
        const uint64_t v[12] = {L.units, L.runs, L.slots, L.subtasks, L.off_deg, L.off_unit, L.off_bin,
                                L.off_rptr, L.off_roff, L.off_rpos, L.off_dst, L.off_src};
        memcpy(info, v, sizeof(v));
    }
    if (graph && prb::build(out_row_ptr, out_col, L, (uint8_t*)graph) != 0) return 0;
    return L.graph_bytes;
}

// Creates the two rank-state address spaces of an instance, zeroed (the binned layout's padding
// slots of the link values are read as zeros and never written).
static bool create_states(Instance* in) {
    for (int i = 0; i < 2; ++i) {
        if (pmCreateMemory(in->state_bytes, &in->state[i]) != pmSuccess) {
            in->state[i] = nullptr;
            return false;
        }
        pmRawMemPtr p;
        if (pmGetRawMemPtr(in->state[i], &p) != pmSuccess) return false;
        memset(p, 0, in->state_bytes);
    }
    return true;
}

// CUs of the device the units are planned for (whole rounds of units per flush).
static uint32_t planning_cus() {
    int cus = 0;
    if (pmHasHipDevice() && uh_device_props(0, nullptr, 0, nullptr, &cus, nullptr, nullptr, nullptr) == 0 && cus > 0)
        return (uint32_t)cus;
    return 256;
}

// out_row_ptr / out_col: out-link CSR (0-based targets), e.g. from ua_webgen.
extern "C" void* ua_pr_create(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                              uint32_t pages_per_subtask, float damping) {
    if (!pages || !out_row_ptr || !pages_per_subtask) {
        apps::set_error("ua_pr_create: invalid arguments");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = prDataDistribution;
        cb.subtask_cpu = pr_cpu;
        cb.subtask_gpu_custom = pr_gpu;
        if (apps::register_variants("PAGERANK", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->pages = pages;
    in->per_subtask = pages_per_subtask;
    in->damping = damping;
    in->edges = out_row_ptr[pages];
    const size_t N = pages;
    in->inputs = std::make_shared<apps::SharedMem>();
    // binned where it can hold the graph (its fixed-point sums assume damping <= 1)
    if (g_layout == LAYOUT_BINNED && damping >= 0.f && damping <= 1.f && in->edges < (1ull << 32)) {
        auto L = std::make_shared<prb::Layout>();
        if (prb::plan(pages, out_row_ptr, out_col, pages_per_subtask, planning_cus(), *L)) {
            in->layout = LAYOUT_BINNED;
            in->binned = L;
            in->state_bytes = L->state_bytes;
            pmRawMemPtr p;
            if (pmCreateMemory(L->graph_bytes, &in->graph) != pmSuccess) in->graph = nullptr;
            in->inputs->handles.push_back(in->graph);
            if (!in->graph || !create_states(in) || pmGetRawMemPtr(in->graph, &p) != pmSuccess) {
                apps::set_error(std::string("create memory: ") + pmGetLastError());
                ua_pr_destroy(in);
                return nullptr;
            }
            if (prb::build(out_row_ptr, out_col, *L, (uint8_t*)p) != 0) {
                apps::set_error("ua_pr_create: binned layout does not add up");
                ua_pr_destroy(in);
                return nullptr;
            }
            return in;
        }
    }
    std::vector<uint8_t> bcnt;
    bool blocked = false;
    if (g_layout != LAYOUT_CSR && in->edges < (1ull << 32)) {
        bcnt.resize((size_t)groups_of(pages) * NB * GR);
        blocked = ua_pr_block_counts(pages, out_row_ptr, out_col, bcnt.data()) == 1;
    }
    in->layout = blocked ? LAYOUT_BLOCKED : LAYOUT_CSR;
    in->state_bytes = 2 * N * 4;
    const size_t graph_bytes = ua_pr_graph_bytes(pages, in->edges, blocked);
    if (pmCreateMemory(graph_bytes, &in->graph) != pmSuccess) in->graph = nullptr;
    in->inputs->handles.push_back(in->graph);
    if (!in->graph || !create_states(in)) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        ua_pr_destroy(in);
        return nullptr;
    }
    pmRawMemPtr p;
    pmGetRawMemPtr(in->graph, &p);
    uint32_t* g = (uint32_t*)p;
    if (ua_pr_build_graph(pages, out_row_ptr, out_col, blocked ? bcnt.data() : nullptr, g) != 0) {
        apps::set_error("ua_pr_create: blocked layout does not add up");
        ua_pr_destroy(in);
        return nullptr;
    }
    in->in_row_ptr = std::make_shared<std::vector<uint32_t>>(g, g + N + 1);
    if (blocked) {
        const uint32_t* bseg = g + 2 * N + 1 + in->edges;
        in->bseg = std::make_shared<std::vector<uint32_t>>(bseg, bseg + (size_t)groups_of(pages) * NB + 1);
    }
    return in;
}

// A second instance over the same graph with its own rank state.
extern "C" void* ua_pr_clone(void* h) {
    const Instance* src = (const Instance*)h;
    if (!src) return nullptr;
    Instance* in = new Instance();
    in->pages = src->pages;
    in->per_subtask = src->per_subtask;
    in->edges = src->edges;
    in->damping = src->damping;
    in->layout = src->layout;
    in->in_row_ptr = src->in_row_ptr;
    in->bseg = src->bseg;
    in->binned = src->binned;
    in->state_bytes = src->state_bytes;
    in->inputs = src->inputs;
    in->graph = src->graph;
    if (!create_states(in)) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        ua_pr_destroy(in);
        return nullptr;
    }
    return in;
}

// Runs the init task and `iterations` iteration tasks on the device set.
extern "C" int ua_pr_run(void* h, int iterations, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2 || iterations < 0) return -1;
    double t = 0, s = 0;
    int rc = run_kind(in, 0, device, &s, iterations > 0);
    t += s;
    for (int i = 0; i < iterations && rc == 0; ++i) {
        rc = run_kind(in, 1, device, &s, i + 1 < iterations);
        t += s;
    }
    if (seconds) *seconds = t;
    return rc;
}

extern "C" int ua_pr_get_ranks(void* h, float* y) {
    Instance* in = (Instance*)h;
    pmMemHandle m = in->state[in->current];
    pmStatus st = pmFetchMemory(m);
    if (st != pmSuccess) return (int)st;
    pmRawMemPtr p;
    st = pmGetRawMemPtr(m, &p);
    if (st != pmSuccess) return (int)st;
    memcpy(y, p, (size_t)in->pages * 4);
    return 0;
}

// Fetches the latest state and hashes its ranks (the first N floats).
extern "C" int ua_pr_fetch(void* h, uint64_t* checksum) {
    Instance* in = (Instance*)h;
    return apps::fetch_output(in->state[in->current], (size_t)in->pages * 4, checksum);
}

extern "C" uint64_t ua_pr_edges(void* h) { return ((Instance*)h)->edges; }

extern "C" int ua_pr_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    if (in->state[0]) pmReleaseMemory(in->state[0]);
    if (in->state[1]) pmReleaseMemory(in->state[1]);
    delete in;
    return 0;
}
