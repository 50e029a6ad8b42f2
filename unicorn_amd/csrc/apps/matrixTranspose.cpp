// matrixTranspose app (reference testSuite/matrixTranspose, square blocks): a rows x cols matrix of
// 32-bit integers is transposed block by block, one subtask per bs x bs block
// (bs = min(rows, cols, 2048), matrixTranspose.cpp:52-61).
//   out of place: input READ_ONLY, output WRITE_ONLY, both COMPACT (matrixTranspose.cpp:474-478);
//   in place: one READ_WRITE space, OPTIMAL visibility; subtasks read the input layout and write the
//   output layout of the same space, so their reads overlap other subtasks' writes and the
//   runtime holds the writes back until the task ends (disjointReadWritesAcrossSubtasks false).
// Data distribution: matrixTransposeDataDistribution (matrixTranspose.cpp:294-322); CPU callback:
// matrixTranspose_cpu (324-378); GPU callback: matrixTranspose_cudaLaunchFunc
// (matrixTransposeCUDA.cu:110-141) with uh_transpose32 in place of the CUDA kernels. The reference
// GPU callback reads an in-place block as if its compact rows were contiguous, which holds for
// square matrices only; here the in-place GPU path locates every read and write row in the
// compact view from the runtime's packing rule (records merged and packed in address order), so
// non-square matrices transpose in place on the GPU as well.
#include "app_common.h"

#include <algorithm>

namespace matrixTranspose {

using namespace pm;

enum { INPLACE_MEM_INDEX = 0 };
enum { INPUT_MEM_INDEX = 0, OUTPUT_MEM_INDEX, MAX_MEM_INDICES };
static constexpr size_t MAX_BLOCK_SIZE = 2048;

// reference matrixTransposeTaskConf (matrixTranspose.h:80-87)
struct matrixTransposeTaskConf {
    size_t matrixDimRows;
    size_t matrixDimCols;
    size_t blockSizeRows;
    size_t blockSizeCols;
    bool inplace;
};

struct Geometry {
    size_t bcr, bcc, bir, bic;     // block counts (rows, cols) and this block's (row, col)
    size_t in_off, out_off;        // byte offsets of the block's first input row / output row
};

static Geometry geometry(const matrixTransposeTaskConf* c, unsigned long sid) {
    Geometry g;
    g.bcr = c->matrixDimRows / c->blockSizeRows;
    g.bcc = c->matrixDimCols / c->blockSizeCols;
    g.bir = sid / g.bcc;
    g.bic = sid % g.bcc;
    const size_t elems = c->blockSizeRows * c->blockSizeCols;
    g.in_off = (g.bic * c->blockSizeCols + g.bir * g.bcc * elems) * 4;
    g.out_off = (g.bir * c->blockSizeRows + g.bic * g.bcr * elems) * 4;
    return g;
}

// In-place GPU tables (in the subtask's scratch buffer, uploaded with it): three arrays of bs
// uint64 for gathering the read rows into a dense block (compact source offset, dense destination
// offset, length), three for scattering the transposed rows to the write rows.
static void compact_row_tables(const matrixTransposeTaskConf* c, const Geometry& g, uint64_t* tab) {
    const size_t bs = c->blockSizeRows, len = bs * 4;
    struct Rec { size_t a, b; };
    std::vector<Rec> rec;
    rec.reserve(2 * bs);
    for (size_t i = 0; i < bs; ++i) rec.push_back({g.in_off + i * c->matrixDimCols * 4, 0});
    for (size_t i = 0; i < bs; ++i) rec.push_back({g.out_off + i * c->matrixDimRows * 4, 0});
    for (auto& r : rec) r.b = r.a + len;
    std::vector<Rec> u(rec);
    std::sort(u.begin(), u.end(), [](const Rec& x, const Rec& y) { return x.a < y.a; });
    std::vector<Rec> merged;                       // union of the records: overlapping or touching merge
    std::vector<size_t> packed;                    // compact offset of each merged range
    size_t total = 0;
    for (const auto& r : u) {
        if (!merged.empty() && r.a <= merged.back().b) {
            merged.back().b = std::max(merged.back().b, r.b);
        } else {
            if (!merged.empty()) total += merged.back().b - merged.back().a;
            merged.push_back(r);
            packed.push_back(total);
        }
    }
    auto compact_of = [&](size_t addr) {
        size_t k = std::upper_bound(merged.begin(), merged.end(), addr, [](size_t v, const Rec& m) { return v < m.a; }) -
                   merged.begin() - 1;
        return packed[k] + (addr - merged[k].a);
    };
    uint64_t* gs = tab;
    uint64_t* gd = tab + bs;
    uint64_t* gl = tab + 2 * bs;
    uint64_t* ss = tab + 3 * bs;
    uint64_t* sd = tab + 4 * bs;
    uint64_t* sl = tab + 5 * bs;
    for (size_t i = 0; i < bs; ++i) {
        gs[i] = compact_of(rec[i].a);
        gd[i] = i * len;
        gl[i] = len;
        ss[i] = i * len;
        sd[i] = compact_of(rec[bs + i].a);
        sl[i] = len;
    }
}

static pmStatus mtDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const matrixTransposeTaskConf* c = (const matrixTransposeTaskConf*)ti.taskConf;
    const Geometry g = geometry(c, si.subtaskId);
    const unsigned in_idx = c->inplace ? INPLACE_MEM_INDEX : INPUT_MEM_INDEX;
    const unsigned out_idx = c->inplace ? INPLACE_MEM_INDEX : OUTPUT_MEM_INDEX;
    pmStatus s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, in_idx, READ_SUBSCRIPTION,
                                     pmScatteredSubscriptionInfo(g.in_off, c->blockSizeCols * 4, c->matrixDimCols * 4,
                                                                 c->blockSizeRows));
    if (s == pmSuccess)
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, out_idx, WRITE_SUBSCRIPTION,
                                pmScatteredSubscriptionInfo(g.out_off, c->blockSizeRows * 4, c->matrixDimRows * 4,
                                                            c->blockSizeCols));
    if (s == pmSuccess && c->inplace && di.deviceType == GPU_HIP) {
        // two dense bs x bs blocks of device memory (reference pmReserveCudaGlobalMem, one block)
        const size_t block = c->blockSizeRows * c->blockSizeRows * 4;
        s = pmReserveGpuGlobalMem(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 2 * block);
        if (s == pmSuccess) {
            uint64_t* tab = (uint64_t*)pmGetScratchBuffer(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo,
                                                          PRE_SUBTASK_TO_SUBTASK, 6 * c->blockSizeRows * 8, NULL);
            if (!tab) return pmMemoryError;
            compact_row_tables(c, g, tab);
        }
    }
    return s;
}

// CPU: matrixTranspose_cpu with square blocks (matrixTranspose.cpp:324-378)
static pmStatus mt_cpu(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const matrixTransposeTaskConf* c = (const matrixTransposeTaskConf*)ti.taskConf;
    const size_t bs = c->blockSizeRows;
    if (c->inplace) {
        const pmMemInfo& m = si.memInfo[INPLACE_MEM_INDEX];
        if (m.visibilityType != SUBSCRIPTION_NATURAL) return pmInvalidCommand;
        const int32_t* src = (const int32_t*)m.readPtr;
        int32_t* dst = (int32_t*)m.writePtr;
        int32_t* blk = (int32_t*)pmGetScratchBuffer(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo,
                                                    PRE_SUBTASK_TO_SUBTASK, bs * bs * 4, NULL);
        if (!src || !dst || !blk) return pmMemoryError;
        for (size_t i = 0; i < bs; ++i) memcpy(blk + i * bs, src + i * c->matrixDimCols, bs * 4);
        for (size_t i = 0; i < bs; ++i)
            for (size_t j = i + 1; j < bs; ++j) std::swap(blk[j + i * bs], blk[i + j * bs]);
        for (size_t i = 0; i < bs; ++i) memcpy(dst + i * c->matrixDimRows, blk + i * bs, bs * 4);
    } else {
        const pmMemInfo& mi = si.memInfo[INPUT_MEM_INDEX];
        const pmMemInfo& mo = si.memInfo[OUTPUT_MEM_INDEX];
        const int32_t* src = (const int32_t*)mi.ptr;
        int32_t* dst = (int32_t*)mo.writePtr;
        const size_t ldi = mi.visibilityType == SUBSCRIPTION_COMPACT ? bs : c->matrixDimCols;
        const size_t ldo = mo.visibilityType == SUBSCRIPTION_COMPACT ? bs : c->matrixDimRows;
        for (size_t i = 0; i < bs; ++i)
            for (size_t j = 0; j < bs; ++j) dst[j + i * ldo] = src[i + j * ldi];
    }
    return pmSuccess;
}

// GPU: the block moves through uh_transpose32 (matrixTransposeCUDA.cu:110-141)
static pmStatus mt_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const matrixTransposeTaskConf* c = (const matrixTransposeTaskConf*)ti.taskConf;
    const size_t bs = c->blockSizeRows;
    int rc;
    if (c->inplace) {
        const pmMemInfo& m = si.memInfo[INPLACE_MEM_INDEX];
        if (m.visibilityType != SUBSCRIPTION_COMPACT || !si.gpuContext.reservedGlobalMem || !si.gpuContext.scratchBuffer)
            return pmInvalidCommand;
        uint8_t* dense = (uint8_t*)si.gpuContext.reservedGlobalMem;
        uint8_t* dense_t = dense + bs * bs * 4;
        const uint64_t* tab = (const uint64_t*)si.gpuContext.scratchBuffer;
        rc = uh_copy_records(dense, m.ptr, tab, tab + bs, tab + 2 * bs, (uint32_t)bs, bs * 4, stream);
        if (!rc) rc = uh_transpose32(dense, bs, dense_t, bs, (uint32_t)bs, (uint32_t)bs, stream);
        if (!rc) rc = uh_copy_records(m.ptr, dense_t, tab + 3 * bs, tab + 4 * bs, tab + 5 * bs, (uint32_t)bs, bs * 4, stream);
    } else {
        const pmMemInfo& mi = si.memInfo[INPUT_MEM_INDEX];
        const pmMemInfo& mo = si.memInfo[OUTPUT_MEM_INDEX];
        const size_t ldi = mi.visibilityType == SUBSCRIPTION_COMPACT ? bs : c->matrixDimCols;
        const size_t ldo = mo.visibilityType == SUBSCRIPTION_COMPACT ? bs : c->matrixDimRows;
        rc = uh_transpose32(mi.ptr, ldi, mo.writePtr, ldo, (uint32_t)bs, (uint32_t)bs, stream);
    }
    if (rc) {
        apps::set_error(std::string("matrixTranspose: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    matrixTransposeTaskConf conf;
    pmMemHandle in = nullptr, out = nullptr;    // in place: one space (in == out)
    unsigned long subtasks = 0;
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

}  // namespace matrixTranspose

using namespace matrixTranspose;

extern "C" void* ua_mt_create(uint64_t rows, uint64_t cols, int inplace) {
    if (!rows || !cols) {
        apps::set_error("ua_mt_create: empty matrix");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = mtDataDistribution;
        cb.subtask_cpu = mt_cpu;
        cb.subtask_gpu_custom = mt_gpu;
        if (apps::register_variants("MATRIXTRANSPOSE", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    const size_t bs = std::min(std::min(rows, cols), (uint64_t)MAX_BLOCK_SIZE);
    if (rows % bs || cols % bs) {
        apps::set_error("ua_mt_create: rows and cols must be multiples of min(rows, cols, 2048)");
        delete in;
        return nullptr;
    }
    in->conf.matrixDimRows = rows;
    in->conf.matrixDimCols = cols;
    in->conf.blockSizeRows = in->conf.blockSizeCols = bs;
    in->conf.inplace = inplace != 0;
    in->subtasks = (unsigned long)((rows / bs) * (cols / bs));
    const size_t bytes = rows * cols * 4;
    if (pmCreateMemory(bytes, &in->in) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    if (in->conf.inplace) in->out = in->in;
    else if (pmCreateMemory(bytes, &in->out) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        pmReleaseMemory(in->in);
        delete in;
        return nullptr;
    }
    return in;
}

extern "C" int ua_mt_set_input(void* h, const int32_t* data) {
    Instance* in = (Instance*)h;
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->in, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, data, in->conf.matrixDimRows * in->conf.matrixDimCols * 4);
    return 0;
}

extern "C" int ua_mt_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    pmTaskMem mem[MAX_MEM_INDICES];
    uint count;
    if (in->conf.inplace) {
        mem[INPLACE_MEM_INDEX] = pmTaskMem(in->out, READ_WRITE, SUBSCRIPTION_OPTIMAL);
        count = 1;
    } else {
        mem[INPUT_MEM_INDEX] = pmTaskMem(in->in, READ_ONLY, SUBSCRIPTION_COMPACT);
        mem[OUTPUT_MEM_INDEX] = pmTaskMem(in->out, WRITE_ONLY, SUBSCRIPTION_COMPACT);
        count = MAX_MEM_INDICES;
    }
    pmTaskDetails d(&in->conf, sizeof(in->conf), mem, count, g_cb.h[device], in->subtasks);
    d.multiAssignEnabled = apps::split_request().multi_assign;
    return apps::run_task(d, seconds);
}

// The transposed matrix (cols x rows).
extern "C" int ua_mt_get_output(void* h, int32_t* out) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->out);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->out, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(out, p, in->conf.matrixDimRows * in->conf.matrixDimCols * 4);
    return 0;
}

extern "C" int ua_mt_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    if (in->out != in->in) pmReleaseMemory(in->out);
    pmReleaseMemory(in->in);
    delete in;
    return 0;
}
