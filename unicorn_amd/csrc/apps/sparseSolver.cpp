// sparseSolver app (reference testSuite/sparseSolver): C = A B with A an n x n sparse float matrix
// in COO form sorted by (row, column), B and C dense n x n, one subtask per bs x bs block of C
// (bs = min(n, 2048); subtask id increases down a block column, sparseSolver.cpp:148-151).
// The task configuration carries, after its two sizes, the first and last COO index of each
// subtask's block row (-1 when the block row has no entries), computed on the host as the reference
// does (sparseSolver.cpp:370-405); the data distribution subscribes to those COO entries, B's block
// column and C's block (sparseSolver.cpp:142-187). The reference app declares six task memories
// (values, B, row indices, column indices, per-row counts, C) while its runtime takes three
// (MAX_MEM_SECTIONS_PER_TASK, pmPublicDefinitions.h:54); here A's entries are one address space of
// 12-byte records {row, column, value}, and the per-row counts, which its CPU callback uses only to
// find the entries of a split piece's rows, are not needed: the piece filters the entries by row.
// CPU callback: sparseMatrixMultiply_cpu (189-263) with the reference's MKL coomm call restated
// as a loop over the entries in COO order (C zeroed first), rows split across pieces as its
// GetSplitData (87-107); GPU callback: uh_spmm_coo_f32 in place of the cuSPARSE coo2csr + csrmm2
// pair (sparseSolverCUDA.cu:206-290). Both accumulate each output in COO order with separately
// rounded products and sums, so CPU and GPU give the same bits.
#include "app_common.h"

#include <algorithm>

namespace sparseSolver {

using namespace pm;

enum { INPUT_ENTRIES_MEM_INDEX = 0, INPUT_MATRIX2_MEM_INDEX, OUTPUT_MATRIX_MEM_INDEX, MAX_MEM_INDICES };

struct CooEntry {
    int32_t row, col;
    float val;
};
static constexpr size_t BLOCK_DIM = 2048;

// reference sparseMatrixMultiplyTaskConf (sparseSolver.h:92-97), followed by 2 ints per subtask
struct sparseMatrixMultiplyTaskConf {
    size_t matrixDim;
    size_t blockDim;
};

static const int* distribution(const sparseMatrixMultiplyTaskConf* c) {
    return (const int*)((const char*)c + sizeof(sparseMatrixMultiplyTaskConf));
}

// rows of the block handled by a split piece (reference GetSplitData, sparseSolver.cpp:87-107)
static bool GetSplitData(size_t* off, size_t* count, const sparseMatrixMultiplyTaskConf* c, const pmSplitInfo& s) {
    *off = 0;
    *count = c->blockDim;
    if (s.splitCount) {
        const size_t n = std::min((size_t)s.splitCount, c->blockDim);
        if (s.splitId > n - 1) return false;
        const size_t per = c->blockDim / n, left = c->blockDim - per * n;
        *count = per + (left > s.splitId ? 1 : 0);
        *off = s.splitId * per + std::min(left, (size_t)s.splitId);
    }
    return true;
}

static pmStatus ssDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const sparseMatrixMultiplyTaskConf* c = (const sparseMatrixMultiplyTaskConf*)ti.taskConf;
    size_t off, count;
    if (!GetSplitData(&off, &count, c, si.splitInfo)) return pmSuccess;
    const size_t bpd = c->matrixDim / c->blockDim;
    const size_t br = si.subtaskId % bpd, bc = si.subtaskId / bpd;
    const int first = distribution(c)[2 * si.subtaskId], last = distribution(c)[2 * si.subtaskId + 1];
    const size_t n = c->matrixDim, bs = c->blockDim, f = sizeof(float);
    pmStatus s = pmSuccess;
    if (first != -1 && last != -1) {
        const size_t k = (size_t)(last - first + 1);
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_ENTRIES_MEM_INDEX,
                                READ_SUBSCRIPTION, pmSubscriptionInfo((size_t)first * sizeof(CooEntry), k * sizeof(CooEntry)));
        if (s == pmSuccess)
            s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MATRIX2_MEM_INDEX,
                                    READ_SUBSCRIPTION, pmScatteredSubscriptionInfo(bc * bs * f, bs * f, n * f, n));
    }
    if (s == pmSuccess)
        s = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_MATRIX_MEM_INDEX,
                                WRITE_SUBSCRIPTION,
                                pmScatteredSubscriptionInfo(((off + br * bs) * n + bc * bs) * f, bs * f, n * f, count));
    return s;
}

static size_t ld_of(const pmMemInfo& m, const sparseMatrixMultiplyTaskConf* c) {
    return m.visibilityType == SUBSCRIPTION_NATURAL ? c->matrixDim : c->blockDim;
}

static pmStatus ss_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const sparseMatrixMultiplyTaskConf* c = (const sparseMatrixMultiplyTaskConf*)ti.taskConf;
    size_t off, count;
    if (!GetSplitData(&off, &count, c, si.splitInfo)) return pmSuccess;
    const size_t bs = c->blockDim, bpd = c->matrixDim / bs, br = si.subtaskId % bpd;
    float* C = (float*)si.memInfo[OUTPUT_MATRIX_MEM_INDEX].ptr;
    const size_t ldc = ld_of(si.memInfo[OUTPUT_MATRIX_MEM_INDEX], c);
    for (size_t i = 0; i < count; ++i) memset(C + i * ldc, 0, bs * sizeof(float));
    const int first = distribution(c)[2 * si.subtaskId], last = distribution(c)[2 * si.subtaskId + 1];
    if (first == -1 || last == -1) return pmSuccess;
    const CooEntry* ent = (const CooEntry*)si.memInfo[INPUT_ENTRIES_MEM_INDEX].ptr;
    const float* B = (const float*)si.memInfo[INPUT_MATRIX2_MEM_INDEX].ptr;
    const size_t ldb = ld_of(si.memInfo[INPUT_MATRIX2_MEM_INDEX], c);
    const int32_t r0 = (int32_t)(br * bs + off), r1 = r0 + (int32_t)count;
    const size_t k = (size_t)(last - first + 1);
    for (size_t e = 0; e < k; ++e) {
        if (ent[e].row < r0 || ent[e].row >= r1) continue;
        const float a = ent[e].val;
        const float* brow = B + (size_t)ent[e].col * ldb;
        float* crow = C + (size_t)(ent[e].row - r0) * ldc;
        for (size_t j = 0; j < bs; ++j) {
            const float p = a * brow[j];
            crow[j] = crow[j] + p;
        }
    }
    return pmSuccess;
}

static pmStatus ss_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const sparseMatrixMultiplyTaskConf* c = (const sparseMatrixMultiplyTaskConf*)ti.taskConf;
    size_t off, count;
    if (!GetSplitData(&off, &count, c, si.splitInfo)) return pmSuccess;
    const size_t bs = c->blockDim, bpd = c->matrixDim / bs, br = si.subtaskId % bpd;
    const int first = distribution(c)[2 * si.subtaskId], last = distribution(c)[2 * si.subtaskId + 1];
    const bool any = first != -1 && last != -1;
    const pmMemInfo& mc = si.memInfo[OUTPUT_MATRIX_MEM_INDEX];
    int rc;
    if (!any) {
        // no entries in this block row: the block is zero
        rc = uh_spmm_coo_f32(nullptr, 0, 0, (uint32_t)count, nullptr, ld_of(mc, c), (float*)mc.ptr, ld_of(mc, c),
                             (uint32_t)bs, stream);
    } else {
        const pmMemInfo& mb = si.memInfo[INPUT_MATRIX2_MEM_INDEX];
        rc = uh_spmm_coo_f32(si.memInfo[INPUT_ENTRIES_MEM_INDEX].ptr, (uint32_t)(last - first + 1),
                             (int32_t)(br * bs + off), (uint32_t)count, (const float*)mb.ptr, ld_of(mb, c),
                             (float*)mc.ptr, ld_of(mc, c), (uint32_t)bs, stream);
    }
    if (rc) {
        apps::set_error(std::string("uh_spmm_coo_f32: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    std::vector<char> conf;
    size_t n = 0, nnz = 0;
    pmMemHandle mem[MAX_MEM_INDICES] = {};
    unsigned long subtasks = 0;
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

}  // namespace sparseSolver

using namespace sparseSolver;

// A: nnz entries (row, col, val) sorted by (row, col), rows and cols < n; B: n x n row-major.
extern "C" void* ua_ss_create(uint64_t n, uint64_t nnz, const int32_t* rows, const int32_t* cols, const float* vals,
                              const float* B) {
    if (!n || (nnz && (!rows || !cols || !vals)) || !B) {
        apps::set_error("ua_ss_create: empty matrix");
        return nullptr;
    }
    const size_t bs = std::min((size_t)n, BLOCK_DIM);
    if (n % bs) {
        apps::set_error("ua_ss_create: n must be a multiple of min(n, 2048)");
        return nullptr;
    }
    for (uint64_t k = 0; k < nnz; ++k) {
        if (rows[k] < 0 || (uint64_t)rows[k] >= n || cols[k] < 0 || (uint64_t)cols[k] >= n ||
            (k && (rows[k] < rows[k - 1] || (rows[k] == rows[k - 1] && cols[k] <= cols[k - 1])))) {
            apps::set_error("ua_ss_create: COO entries must be in range and strictly sorted by (row, col)");
            return nullptr;
        }
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = ssDataDistribution;
        cb.subtask_cpu = ss_cpu;
        cb.subtask_gpu_custom = ss_gpu;
        if (apps::register_variants("SPARSEMATRIXMUL", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->n = n;
    in->nnz = nnz;
    const size_t bpd = n / bs;
    in->subtasks = (unsigned long)(bpd * bpd);
    in->conf.assign(sizeof(sparseMatrixMultiplyTaskConf) + 2 * sizeof(int) * in->subtasks, 0);
    sparseMatrixMultiplyTaskConf* c = (sparseMatrixMultiplyTaskConf*)in->conf.data();
    c->matrixDim = n;
    c->blockDim = bs;
    // first and last entry of every subtask's block row (reference sparseSolver.cpp:370-405)
    int* dist = (int*)(in->conf.data() + sizeof(sparseMatrixMultiplyTaskConf));
    for (unsigned long s = 0; s < in->subtasks; ++s) {
        const int64_t r0 = (int64_t)(s % bpd) * bs, r1 = r0 + (int64_t)bs;
        const int32_t* lo = std::lower_bound(rows, rows + nnz, (int32_t)r0);
        const int32_t* hi = std::lower_bound(rows, rows + nnz, (int32_t)std::min<int64_t>(r1, INT32_MAX));
        if (lo == hi) dist[2 * s] = dist[2 * s + 1] = -1;
        else {
            dist[2 * s] = (int)(lo - rows);
            dist[2 * s + 1] = (int)(hi - rows) - 1;
        }
    }
    const size_t sizes[MAX_MEM_INDICES] = {std::max<size_t>(nnz, 1) * sizeof(CooEntry), n * n * 4, n * n * 4};
    for (int i = 0; i < MAX_MEM_INDICES; ++i)
        if (pmCreateMemory(sizes[i], &in->mem[i]) != pmSuccess) {
            apps::set_error(std::string("create memory: ") + pmGetLastError());
            for (int j = 0; j < i; ++j) pmReleaseMemory(in->mem[j]);
            delete in;
            return nullptr;
        }
    pmRawMemPtr p;
    pmGetRawMemPtr(in->mem[INPUT_ENTRIES_MEM_INDEX], &p);
    for (uint64_t k = 0; k < nnz; ++k) ((CooEntry*)p)[k] = CooEntry{rows[k], cols[k], vals[k]};
    pmGetRawMemPtr(in->mem[INPUT_MATRIX2_MEM_INDEX], &p);
    memcpy(p, B, n * n * 4);
    return in;
}

extern "C" int ua_ss_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    pmTaskMem mem[MAX_MEM_INDICES];
    for (int i = 0; i < MAX_MEM_INDICES; ++i)
        mem[i] = pmTaskMem(in->mem[i], i == OUTPUT_MATRIX_MEM_INDEX ? WRITE_ONLY : READ_ONLY, SUBSCRIPTION_OPTIMAL);
    pmTaskDetails d(in->conf.data(), (uint)in->conf.size(), mem, MAX_MEM_INDICES, g_cb.h[device], in->subtasks);
    d.canSplitCpuSubtasks = apps::split_request().cpu;     // as the reference (sparseSolver.cpp:366)
    d.canSplitGpuSubtasks = apps::split_request().gpu;
    d.multiAssignEnabled = apps::split_request().multi_assign;
    return apps::run_task(d, seconds);
}

extern "C" int ua_ss_get_output(void* h, float* C) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->mem[OUTPUT_MATRIX_MEM_INDEX]);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->mem[OUTPUT_MATRIX_MEM_INDEX], &p);
    if (s != pmSuccess) return (int)s;
    memcpy(C, p, in->n * in->n * 4);
    return 0;
}

extern "C" int ua_ss_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    for (int i = 0; i < MAX_MEM_INDICES; ++i) pmReleaseMemory(in->mem[i]);
    delete in;
    return 0;
}
