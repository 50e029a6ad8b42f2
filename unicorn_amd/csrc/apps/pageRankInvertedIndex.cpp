// pageRankInvertedIndex app (reference testSuite/pageRankInvertedIndex): pageRank over the
// reference's on-disk web dump in two steps (DoParallelProcess, pageRankInvertedIndex.cpp:639-700):
//   1. the invert task: subtask s reads the mapped dump files of its pages (LoadMappedFiles,
//      249-277) and writes one inverted-index entry {src page, dest page, src out-links} per link
//      into its window of the index space, grouped by the subtask that owns the destination page
//      (invertIndex_cpu, 301-382: counts per group into the SUBTASK_TO_POST_SUBTASK scratch buffer,
//      then each group filled from its highest position down); its dataRedistribution callback
//      (384-399) hands group g to order g, so the runtime's redistribution leaves every group's
//      entries from all subtasks together, in order (pmRedistributeData);
//   2. PAGE_RANK_ITERATIONS (2) tasks: subtask g reads its inlinks (the redistribution metadata
//      of order g), all of the previous ranks, and writes its pages' ranks:
//      out[dest] += 0.85 * (x[src] / out-links), x = 1.0 on the first iteration, no teleport term
//      (pageRankDataDistribution / pageRank_cpu, 401-472).
// Differences from the reference, all where it would fail:
//   - a subtask's first page is s * P (the reference computes it from the subtask's own file count,
//     which is wrong for a shorter last subtask);
//   - a group that receives no inlinks gets no redistribution order (as in the reference's runtime)
//     and its pageRank subtask zero inlinks (the reference exits on the missing order, 573-577);
//   - the scratch buffer is requested by the data distribution callback, so that the HIP stub
//     stages it (the reference takes it inside invertIndex_cpu);
//   - malformed records (count > max out-links, a target outside 1..pages) fail the subtask.
// GPU callbacks: the reference's are empty (pageRankInvertedIndexCUDA.cu:92-100). Here the invert
// callback copies its pages' records from the mapped files to HBM and runs uh_invert_index (same
// entries, same order); the pageRank callback runs uh_prii_rank (float atomics: within rounding of
// the CPU callback's entry-order sums).
#include "app_common.h"

#include <algorithm>
#include <string>

namespace pageRankInvertedIndex {

using namespace pm;

static constexpr float DAMPENING_FACTOR = 0.85f;
static constexpr float INITIAL_PAGE_RANK = 1.0f;
static constexpr unsigned int PAGE_RANK_ITERATIONS = 2;
static constexpr unsigned int WEB_PAGES_PER_SUBTASK = 10000000;
static constexpr size_t MAX_BASE_PATH_LENGTH = 256;

// reference pageRankInvertedIndex.h:53-85
struct invertedIndex {
    unsigned int srcPage;
    unsigned int destPage;          // srcPage has an outlink to destPage (0-based)
    unsigned int srcPageOutlinks;
};
struct subtaskMetadata {
    unsigned long inlinksStartOffset;
    unsigned int inlinksCount;
};
struct invertIndexTaskConf {
    unsigned int totalWebPages;
    unsigned int maxOutlinksPerWebPage;
    unsigned int webPagesPerFile;
    unsigned int webPagesPerSubtask;
    char basePath[MAX_BASE_PATH_LENGTH];
};
struct pageRankTaskConf {
    unsigned int totalWebPages;
    unsigned int maxOutlinksPerWebPage;
    unsigned int webPagesPerFile;
    unsigned int webPagesPerSubtask;
    unsigned int iteration;
    float initialPageRank;
    subtaskMetadata* metadata;      // unused: the entries follow the struct in the task conf
};

enum memIndex { INVERTED_INDEX_MEM_INDEX = 0 };
enum memIndex2 { INVERTED_LINKS_MEM_INDEX = 0, OUTPUT_PAGE_RANKS_MEM_INDEX = 1, INPUT_PAGE_RANKS_MEM_INDEX = 2 };

template <typename C>
static void page_range(const C* c, ulong sid, unsigned int* first, unsigned int* count) {
    const uint64_t f = (uint64_t)sid * c->webPagesPerSubtask;
    *first = (unsigned int)f;
    *count = (unsigned int)std::min<uint64_t>(c->webPagesPerSubtask, c->totalWebPages - f);
}

static std::string page_file(const char* dir, uint64_t first_page0) {
    return std::string(dir) + "/web/page_" + std::to_string(first_page0 + 1);
}

// Mapped records of the subtask's files, one pointer per file (LoadMappedFiles).
static bool mapped_files(const invertIndexTaskConf* c, unsigned int first, unsigned int count,
                         std::vector<const uint32_t*>& files) {
    files.clear();
    for (uint64_t p = first; p < (uint64_t)first + count; p += c->webPagesPerFile) {
        const uint32_t* f = (const uint32_t*)pmGetMappedFile(page_file(c->basePath, p).c_str());
        if (!f) return false;
        files.push_back(f);
    }
    return true;
}

static pmStatus invertIndexDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const invertIndexTaskConf* c = (const invertIndexTaskConf*)ti.taskConf;
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    const size_t rec = (size_t)c->maxOutlinksPerWebPage * sizeof(invertedIndex);
    if (rec && count) {
        pmStatus s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo,
                                         INVERTED_INDEX_MEM_INDEX, WRITE_SUBSCRIPTION,
                                         pmSubscriptionInfo((size_t)first * rec, (size_t)count * rec));
        if (s != pmSuccess) return s;
    }
    // entries per destination group, read by the redistribution callback
    if (!pmGetScratchBuffer(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, SUBTASK_TO_POST_SUBTASK,
                            ti.subtaskCount * sizeof(unsigned int), nullptr))
        return pmMemoryError;
    return pmSuccess;
}

static pmStatus invertIndex_cpu(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const invertIndexTaskConf* c = (const invertIndexTaskConf*)ti.taskConf;
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    std::vector<const uint32_t*> files;
    if (!mapped_files(c, first, count, files)) return pmInvalidKey;
    unsigned int* perGroup = (unsigned int*)pmGetScratchBuffer(ti.taskHandle, di.deviceHandle, si.subtaskId,
                                                               si.splitInfo, SUBTASK_TO_POST_SUBTASK,
                                                               ti.subtaskCount * sizeof(unsigned int), nullptr);
    if (!perGroup) return pmMemoryError;
    memset(perGroup, 0, ti.subtaskCount * sizeof(unsigned int));
    const size_t recLen = (size_t)c->maxOutlinksPerWebPage + 1;
    // count the links per destination group
    for (size_t f = 0; f < files.size(); ++f) {
        const unsigned int pages = std::min<unsigned int>(c->webPagesPerFile, count - (unsigned int)f * c->webPagesPerFile);
        for (unsigned int j = 0; j < pages; ++j) {
            const uint32_t* r = files[f] + j * recLen;
            if (r[0] > c->maxOutlinksPerWebPage) return pmDataProcessingFailure;
            for (uint32_t k = 0; k < r[0]; ++k) {
                if (r[1 + k] < 1 || r[1 + k] > c->totalWebPages) return pmDataProcessingFailure;
                ++perGroup[(r[1 + k] - 1) / c->webPagesPerSubtask];
            }
        }
    }
    // highest position of each group, then fill downwards in page / link order
    std::vector<size_t> loc(ti.subtaskCount);
    size_t end = 0;
    for (ulong g = 0; g < ti.subtaskCount; ++g) {
        end += perGroup[g];
        loc[g] = end;          // one past the group's highest position
    }
    invertedIndex* out = (invertedIndex*)si.memInfo[INVERTED_INDEX_MEM_INDEX].ptr;
    for (size_t f = 0; f < files.size(); ++f) {
        const unsigned int pages = std::min<unsigned int>(c->webPagesPerFile, count - (unsigned int)f * c->webPagesPerFile);
        for (unsigned int j = 0; j < pages; ++j) {
            const uint32_t* r = files[f] + j * recLen;
            const unsigned int page = first + (unsigned int)f * c->webPagesPerFile + j;
            for (uint32_t k = 0; k < r[0]; ++k) {
                invertedIndex& e = out[--loc[(r[1 + k] - 1) / c->webPagesPerSubtask]];
                e.srcPage = page;
                e.destPage = r[1 + k] - 1;
                e.srcPageOutlinks = r[0];
            }
        }
    }
    return pmSuccess;
}

static pmStatus invertIndex_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const invertIndexTaskConf* c = (const invertIndexTaskConf*)ti.taskConf;
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    std::vector<const uint32_t*> files;
    if (!mapped_files(c, first, count, files)) return pmInvalidKey;
    uint32_t* totals = (uint32_t*)si.gpuContext.scratchBuffer;
    if (!totals) return pmMemoryError;
    const size_t recBytes = ((size_t)c->maxOutlinksPerWebPage + 1) * 4;
    void* dump = nullptr;
    int rc = uh_malloc(&dump, (size_t)count * recBytes);
    for (size_t f = 0; rc == 0 && f < files.size(); ++f) {
        const unsigned int pages = std::min<unsigned int>(c->webPagesPerFile, count - (unsigned int)f * c->webPagesPerFile);
        rc = uh_memcpy_h2d((uint8_t*)dump + f * c->webPagesPerFile * recBytes, files[f], (size_t)pages * recBytes, stream);
    }
    if (rc == 0)
        rc = uh_invert_index((const uint32_t*)dump, count, c->maxOutlinksPerWebPage, first, c->totalWebPages,
                             c->webPagesPerSubtask, (uint32_t)ti.subtaskCount, si.memInfo[INVERTED_INDEX_MEM_INDEX].ptr,
                             totals, stream);
    const std::string err = rc ? std::string("uh_invert_index: ") + uh_last_error() : std::string();
    uh_free(dump);
    if (rc) {
        apps::set_error(err);
        return pmDataProcessingFailure;
    }
    return pmSuccess;
}

static pmStatus invertIndexDataRedistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const unsigned int* perGroup = (const unsigned int*)pmGetScratchBuffer(
        ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, SUBTASK_TO_POST_SUBTASK, 0, nullptr);
    if (!perGroup) return pmMemoryError;
    size_t offset = 0;
    for (ulong g = 0; g < ti.subtaskCount; ++g) {
        const size_t len = (size_t)perGroup[g] * sizeof(invertedIndex);
        pmStatus s = pmRedistributeData(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo,
                                        INVERTED_INDEX_MEM_INDEX, offset, len, (uint)g);
        if (s != pmSuccess) return s;
        offset += len;
    }
    return pmSuccess;
}

static const subtaskMetadata& metadata_of(const pageRankTaskConf* c, ulong sid) {
    return ((const subtaskMetadata*)((const char*)c + sizeof(pageRankTaskConf)))[sid];
}

static pmStatus pageRankDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    const subtaskMetadata& m = metadata_of(c, si.subtaskId);
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    pmStatus s = pmSuccess;
    if (m.inlinksCount)
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INVERTED_LINKS_MEM_INDEX,
                                READ_SUBSCRIPTION,
                                pmSubscriptionInfo(m.inlinksStartOffset, (size_t)m.inlinksCount * sizeof(invertedIndex)));
    if (s == pmSuccess)
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_PAGE_RANKS_MEM_INDEX,
                                WRITE_SUBSCRIPTION, pmSubscriptionInfo((size_t)first * 4, (size_t)count * 4));
    if (s == pmSuccess && c->iteration != 0)
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_PAGE_RANKS_MEM_INDEX,
                                READ_SUBSCRIPTION, pmSubscriptionInfo(0, (size_t)c->totalWebPages * 4));
    return s;
}

static pmStatus pageRank_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    const subtaskMetadata& m = metadata_of(c, si.subtaskId);
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    float* out = (float*)si.memInfo[OUTPUT_PAGE_RANKS_MEM_INDEX].ptr;
    const invertedIndex* in = (const invertedIndex*)si.memInfo[INVERTED_LINKS_MEM_INDEX].readPtr;
    memset(out, 0, (size_t)count * 4);
    if (c->iteration == 0) {
        for (unsigned int i = 0; i < m.inlinksCount; ++i)
            out[in[i].destPage - first] += DAMPENING_FACTOR * (c->initialPageRank / in[i].srcPageOutlinks);
    } else {
        const float* x = (const float*)si.memInfo[INPUT_PAGE_RANKS_MEM_INDEX].ptr;
        for (unsigned int i = 0; i < m.inlinksCount; ++i)
            out[in[i].destPage - first] += DAMPENING_FACTOR * (x[in[i].srcPage] / in[i].srcPageOutlinks);
    }
    return pmSuccess;
}

static pmStatus pageRank_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const pageRankTaskConf* c = (const pageRankTaskConf*)ti.taskConf;
    const subtaskMetadata& m = metadata_of(c, si.subtaskId);
    unsigned int first, count;
    page_range(c, si.subtaskId, &first, &count);
    const float* x = c->iteration ? (const float*)si.memInfo[INPUT_PAGE_RANKS_MEM_INDEX].ptr : nullptr;
    if (uh_prii_rank(si.memInfo[INVERTED_LINKS_MEM_INDEX].readPtr, m.inlinksCount, x, c->initialPageRank,
                     DAMPENING_FACTOR, first, (float*)si.memInfo[OUTPUT_PAGE_RANKS_MEM_INDEX].ptr, count,
                     c->totalWebPages, stream) != 0) {
        apps::set_error(std::string("uh_prii_rank: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    invertIndexTaskConf conf;
    unsigned long subtasks = 0;
    std::vector<std::string> files;
    pmMemHandle index = nullptr, ranks = nullptr;
    std::vector<unsigned int> inlinks;     // per pageRank subtask, from the redistribution metadata
};

static apps::CallbackVariants g_cb_invert, g_cb_rank;
static std::mutex g_mtx;

static int fail(const std::string& m) {
    apps::set_error(m);
    return -1;
}

}  // namespace pageRankInvertedIndex

using namespace pageRankInvertedIndex;

// pages_per_subtask: 0 for the reference's 10,000,000; rounded to whole files as DoInit does
// (pageRankInvertedIndex.cpp:737-760).
extern "C" void* ua_prii_create(const char* dir, uint32_t pages_per_subtask) {
    uint32_t pages = 0, max_out = 0, per_file = 0;
    if (!dir || strlen(dir) >= MAX_BASE_PATH_LENGTH) {
        fail("ua_prii_create: missing or too long base path");
        return nullptr;
    }
    if (ua_webdump_meta(dir, &pages, &max_out, &per_file) != 0) return nullptr;
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks a, b;
        a.dataDistribution = invertIndexDataDistribution;
        a.subtask_cpu = invertIndex_cpu;
        a.subtask_gpu_custom = invertIndex_gpu;
        a.dataRedistribution = invertIndexDataRedistribution;
        b.dataDistribution = pageRankDataDistribution;
        b.subtask_cpu = pageRank_cpu;
        b.subtask_gpu_custom = pageRank_gpu;
        if (apps::register_variants("INVERTINDEX", a, g_cb_invert) != 0 ||
            apps::register_variants("PAGERANK_II", b, g_cb_rank) != 0) {
            fail(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    const uint32_t want = pages_per_subtask ? pages_per_subtask : WEB_PAGES_PER_SUBTASK;
    uint32_t P;
    if (per_file > want) P = per_file;
    else if (want % per_file == 0) P = want;
    else P = (want / per_file) * per_file;
    Instance* in = new Instance();
    memset(&in->conf, 0, sizeof(in->conf));
    in->conf.totalWebPages = pages;
    in->conf.maxOutlinksPerWebPage = max_out;
    in->conf.webPagesPerFile = per_file;
    in->conf.webPagesPerSubtask = P;
    strcpy(in->conf.basePath, dir);
    in->subtasks = (pages + (uint64_t)P - 1) / P;
    for (uint64_t p = 0; p < pages; p += per_file) in->files.push_back(page_file(dir, p));
    std::vector<const char*> paths;
    for (auto& f : in->files) paths.push_back(f.c_str());
    if (pmMapFiles(paths.data(), (uint)paths.size()) != pmSuccess) {
        fail(std::string("map web dump files: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    const size_t index_bytes = std::max<size_t>(1, (size_t)pages * max_out * sizeof(invertedIndex));
    if (pmCreateMemory(index_bytes, &in->index) != pmSuccess) {
        fail(std::string("create memory: ") + pmGetLastError());
        pmUnmapFiles(paths.data(), (uint)paths.size());
        delete in;
        return nullptr;
    }
    return in;
}

// The invert task, then `iterations` pageRank tasks (0: the reference's 2).
extern "C" int ua_prii_run(void* h, int device, uint32_t iterations, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    if (!iterations) iterations = PAGE_RANK_ITERATIONS;
    const double t0 = apps::now_secs();
    pmTaskMem mem[1] = {pmTaskMem(in->index, WRITE_ONLY)};
    pmTaskDetails d(&in->conf, sizeof(in->conf), mem, 1, g_cb_invert.h[device], in->subtasks);
    d.multiAssignEnabled = apps::split_request().multi_assign;
    pmTaskHandle th = nullptr;
    pmStatus s = pmSubmitTask(d, &th);
    if (s != pmSuccess) return (int)s;
    s = pmWaitForTaskCompletion(th);
    if (s != pmSuccess) {
        pmReleaseTask(th);
        return (int)s;
    }
    // inlinks of every pageRank subtask from the redistribution metadata (InvertIndexTask, 546-599)
    ulong orders = 0;
    const pmRedistributionMetadata* meta = pmGetRedistributionMetadata(th, INVERTED_INDEX_MEM_INDEX, &orders);
    std::vector<char> confBuf(sizeof(pageRankTaskConf) + sizeof(subtaskMetadata) * in->subtasks, 0);
    pageRankTaskConf* pc = (pageRankTaskConf*)confBuf.data();
    subtaskMetadata* sm = (subtaskMetadata*)(confBuf.data() + sizeof(pageRankTaskConf));
    in->inlinks.assign(in->subtasks, 0);
    for (ulong k = 0; k < orders; ++k) {
        if (meta[k].order >= in->subtasks || meta[k].count % sizeof(invertedIndex)) {
            pmReleaseTask(th);
            return fail("inverted index: unexpected redistribution order");
        }
        in->inlinks[meta[k].order] = meta[k].count / sizeof(invertedIndex);
    }
    pmReleaseTask(th);
    unsigned long offset = 0;
    for (ulong g = 0; g < in->subtasks; ++g) {
        sm[g].inlinksStartOffset = offset;
        sm[g].inlinksCount = in->inlinks[g];
        offset += (unsigned long)in->inlinks[g] * sizeof(invertedIndex);
    }
    pc->totalWebPages = in->conf.totalWebPages;
    pc->maxOutlinksPerWebPage = in->conf.maxOutlinksPerWebPage;
    pc->webPagesPerFile = in->conf.webPagesPerFile;
    pc->webPagesPerSubtask = in->conf.webPagesPerSubtask;
    pc->initialPageRank = INITIAL_PAGE_RANK;
    // ParallelPageRankIteration (601-637): a fresh output space per iteration
    for (uint32_t it = 0; it < iterations; ++it) {
        pc->iteration = it;
        pmMemHandle out = nullptr;
        if (pmCreateMemory((size_t)in->conf.totalWebPages * 4, &out) != pmSuccess) return (int)pmMemoryError;
        pmTaskMem m2[3] = {pmTaskMem(in->index, READ_ONLY), pmTaskMem(out, WRITE_ONLY), pmTaskMem(in->ranks, READ_ONLY)};
        pmTaskDetails d2(confBuf.data(), (uint)confBuf.size(), m2, in->ranks ? 3 : 2, g_cb_rank.h[device], in->subtasks);
        d2.multiAssignEnabled = apps::split_request().multi_assign;
        const int rc = apps::run_task(d2, nullptr);
        if (rc) {
            pmReleaseMemory(out);
            return rc;
        }
        if (in->ranks) pmReleaseMemory(in->ranks);
        in->ranks = out;
    }
    if (seconds) *seconds = apps::now_secs() - t0;
    return 0;
}

// The redistributed inverted index: n_entries (sum of the inlink counts) entries of three
// uint32 into entries (capacity in entries), and the inlink count of every pageRank subtask.
extern "C" int ua_prii_get_index(void* h, uint32_t* entries, uint64_t capacity, uint64_t* n_entries, uint32_t* counts) {
    Instance* in = (Instance*)h;
    uint64_t n = 0;
    for (unsigned int c : in->inlinks) n += c;
    if (n_entries) *n_entries = n;
    if (counts)
        for (size_t g = 0; g < in->inlinks.size(); ++g) counts[g] = in->inlinks[g];
    if (!entries) return 0;
    if (capacity < n) return fail("ua_prii_get_index: capacity below the entry count");
    pmStatus s = pmFetchMemory(in->index);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->index, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(entries, p, n * sizeof(invertedIndex));
    return 0;
}

extern "C" int ua_prii_get_ranks(void* h, float* ranks) {
    Instance* in = (Instance*)h;
    if (!in->ranks) return fail("ua_prii_get_ranks: no ranks yet");
    pmStatus s = pmFetchMemory(in->ranks);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->ranks, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(ranks, p, (size_t)in->conf.totalWebPages * 4);
    return 0;
}

extern "C" int ua_prii_info(void* h, uint32_t* pages, uint32_t* pages_per_subtask, uint32_t* subtasks) {
    Instance* in = (Instance*)h;
    if (pages) *pages = in->conf.totalWebPages;
    if (pages_per_subtask) *pages_per_subtask = in->conf.webPagesPerSubtask;
    if (subtasks) *subtasks = (uint32_t)in->subtasks;
    return 0;
}

extern "C" int ua_prii_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    if (in->ranks) pmReleaseMemory(in->ranks);
    pmReleaseMemory(in->index);
    std::vector<const char*> paths;
    for (auto& f : in->files) paths.push_back(f.c_str());
    pmUnmapFiles(paths.data(), (uint)paths.size());
    delete in;
    return 0;
}
