// matrixMultiply app on the pm:: API: the block decomposition and subscriptions of the
// reference's matrixMultiplyBlas (testSuite/matrixMultiplyBlas/source/code/src/
// matrixMultiplyBlas.cpp:67-133, matrixMultiplyBlas.h:56-63) with the "plain" i-j-k CPU method of
// testSuite/matrixMultiply (matrixMultiply.cpp:47-55) in float, and a split-BF16 MFMA GPU
// callback in place of cublasSgemm (matrixMultiplyBlasCUDA.cu:54-77).
//
// Subtask s computes output block (s % bpd, s / bpd) ("vertical" order: consecutive subtasks
// share a B column block). It reads the A row block, the B column block, writes the C block.
#include "app_common.h"
#include "unicorn_hip.h"

namespace matrixMultiply {

using namespace pm;

enum memIndex { INPUT_MATRIX1_MEM_INDEX = 0, INPUT_MATRIX2_MEM_INDEX, OUTPUT_MATRIX_MEM_INDEX, MAX_MEM_INDICES };

struct matrixMultiplyTaskConf {
    size_t matrixDim;
    size_t blockDim;
    int precision;     // uh_gemm precision of the GPU callback
};

// Rows of the block handled by a split subtask (reference GetSplitData, matrixMultiplyBlas.cpp:67-88).
static bool GetSplitData(size_t* off, size_t* height, const matrixMultiplyTaskConf* c, const pmSplitInfo& s) {
    *off = 0;
    *height = c->blockDim;
    if (s.splitCount) {
        const size_t count = c->blockDim < s.splitCount ? c->blockDim : s.splitCount;
        if (s.splitId > count - 1) return false;
        const size_t part = c->blockDim / count;
        *off = s.splitId * part;
        *height = (s.splitId == count - 1) ? c->blockDim - s.splitId * part : part;
    }
    return true;
}

static pmStatus mmDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const matrixMultiplyTaskConf* c = (const matrixMultiplyTaskConf*)ti.taskConf;
    const size_t bpd = c->matrixDim / c->blockDim;
    const size_t br = si.subtaskId % bpd, bc = si.subtaskId / bpd;
    size_t off, h;
    if (!GetSplitData(&off, &h, c, si.splitInfo)) return pmSuccess;
    const size_t n = c->matrixDim, bs = c->blockDim, e = sizeof(float);
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MATRIX1_MEM_INDEX, READ_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo((br * bs + off) * n * e, n * e, n * e, h));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MATRIX2_MEM_INDEX, READ_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo(bc * bs * e, bs * e, n * e, n));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_MATRIX_MEM_INDEX, WRITE_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo(((br * bs + off) * n + bc * bs) * e, bs * e, n * e, h));
    return pmSuccess;
}

// CPU: plain i-j-k loop in float over the block, C zeroed first (matrixMultiply.cpp:47-55,110).
static pmStatus mm_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const matrixMultiplyTaskConf* c = (const matrixMultiplyTaskConf*)ti.taskConf;
    size_t off, h;
    if (!GetSplitData(&off, &h, c, si.splitInfo)) return pmSuccess;
    const float* A = (const float*)si.memInfo[INPUT_MATRIX1_MEM_INDEX].ptr;
    const float* B = (const float*)si.memInfo[INPUT_MATRIX2_MEM_INDEX].ptr;
    float* C = (float*)si.memInfo[OUTPUT_MATRIX_MEM_INDEX].ptr;
    const size_t n = c->matrixDim, bs = c->blockDim;
    const size_t ldb = si.memInfo[INPUT_MATRIX2_MEM_INDEX].visibilityType == SUBSCRIPTION_NATURAL ? n : bs;
    const size_t ldc = si.memInfo[OUTPUT_MATRIX_MEM_INDEX].visibilityType == SUBSCRIPTION_NATURAL ? n : bs;
    for (size_t i = 0; i < h; ++i)
        for (size_t j = 0; j < bs; ++j) {
            float s = 0.f;
            for (size_t k = 0; k < n; ++k) s += A[i * n + k] * B[k * ldb + j];
            C[i * ldc + j] = s;
        }
    return pmSuccess;
}

// GPU: C_blk = A_rows * B_cols, row-major (the reference passes the transposed problem to
// column-major cuBLAS; uh_gemm is row-major).
static pmStatus mm_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const matrixMultiplyTaskConf* c = (const matrixMultiplyTaskConf*)ti.taskConf;
    size_t off, h;
    if (!GetSplitData(&off, &h, c, si.splitInfo)) return pmSuccess;
    const size_t n = c->matrixDim, bs = c->blockDim;
    uh_gemm_args g;
    g.A = (const float*)si.memInfo[INPUT_MATRIX1_MEM_INDEX].ptr;
    g.B = (const float*)si.memInfo[INPUT_MATRIX2_MEM_INDEX].ptr;
    g.C = (float*)si.memInfo[OUTPUT_MATRIX_MEM_INDEX].ptr;
    g.lda = n;
    g.ldb = si.memInfo[INPUT_MATRIX2_MEM_INDEX].visibilityType == SUBSCRIPTION_NATURAL ? n : bs;
    g.ldc = si.memInfo[OUTPUT_MATRIX_MEM_INDEX].visibilityType == SUBSCRIPTION_NATURAL ? n : bs;
    g.M = (int)h;
    g.N = (int)bs;
    g.K = (int)n;
    g.precision = c->precision;
    g.alpha = 1.f;
    g.beta = 0.f;
    if (uh_gemm(&g, stream) != 0) {
        apps::set_error(std::string("uh_gemm: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    matrixMultiplyTaskConf conf;
    std::shared_ptr<apps::SharedMem> inputs;   // A and B, shared with clones
    pmMemHandle a = nullptr, b = nullptr, c = nullptr;
    unsigned long subtasks = 0;
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

}  // namespace matrixMultiply

using namespace matrixMultiply;

extern "C" void* ua_mm_create(int n, int block, int precision) {
    if (n <= 0 || block <= 0 || n % block != 0 || precision < 0 || precision > 2) {
        apps::set_error("ua_mm_create: n must be a positive multiple of block");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = mmDataDistribution;
        cb.subtask_cpu = mm_cpu;
        cb.subtask_gpu_custom = mm_gpu;
        if (apps::register_variants("MATRIXMUL", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->conf.matrixDim = (size_t)n;
    in->conf.blockDim = (size_t)block;
    in->conf.precision = precision;
    in->subtasks = (unsigned long)(n / block) * (n / block);
    const size_t row = (size_t)n * sizeof(float);
    in->inputs = std::make_shared<apps::SharedMem>();
    if (pmCreateMemory2D(n, row, &in->a) != pmSuccess) in->a = nullptr;
    in->inputs->handles.push_back(in->a);
    if (in->a && pmCreateMemory2D(n, row, &in->b) != pmSuccess) in->b = nullptr;
    in->inputs->handles.push_back(in->b);
    if (!in->a || !in->b || pmCreateMemory2D(n, row, &in->c) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    // The (n / block)^2 products of a task run in one batched flush, which fills the device the
    // way one n x n x n product does: tune on that, record for the subtask's product shape.
    if (n / block >= 2) apps::autotune_gemm_as(n, n, n, block, block, n, precision);
    else apps::autotune_gemm(block, block, n, precision);
    return in;
}

// A second instance over the same A and B with its own C (for concurrent tasks on one input).
extern "C" void* ua_mm_clone(void* h) {
    const Instance* src = (const Instance*)h;
    if (!src) return nullptr;
    Instance* in = new Instance();
    in->conf = src->conf;
    in->subtasks = src->subtasks;
    in->inputs = src->inputs;
    in->a = src->a;
    in->b = src->b;
    const size_t n = src->conf.matrixDim;
    if (pmCreateMemory2D(n, n * sizeof(float), &in->c) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    return in;
}

extern "C" int ua_mm_set_inputs(void* h, const float* A, const float* B) {
    Instance* in = (Instance*)h;
    const size_t bytes = in->conf.matrixDim * in->conf.matrixDim * sizeof(float);
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->a, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, A, bytes);
    s = pmGetRawMemPtr(in->b, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, B, bytes);
    return 0;
}

extern "C" int ua_mm_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    pmTaskMem mem[MAX_MEM_INDICES];
    mem[INPUT_MATRIX1_MEM_INDEX] = pmTaskMem(in->a, READ_ONLY, SUBSCRIPTION_OPTIMAL);
    mem[INPUT_MATRIX2_MEM_INDEX] = pmTaskMem(in->b, READ_ONLY, SUBSCRIPTION_OPTIMAL);
    mem[OUTPUT_MATRIX_MEM_INDEX] = pmTaskMem(in->c, WRITE_ONLY, SUBSCRIPTION_OPTIMAL);
    pmTaskDetails d(&in->conf, sizeof(in->conf), mem, MAX_MEM_INDICES, g_cb.h[device], in->subtasks);
    d.canSplitCpuSubtasks = apps::split_request().cpu;
    d.canSplitGpuSubtasks = apps::split_request().gpu;
    d.multiAssignEnabled = apps::split_request().multi_assign;
    return apps::run_task(d, seconds);
}

extern "C" int ua_mm_get_output(void* h, float* C) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->c);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->c, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(C, p, in->conf.matrixDim * in->conf.matrixDim * sizeof(float));
    return 0;
}

extern "C" int ua_mm_fetch(void* h, uint64_t* checksum) {
    Instance* in = (Instance*)h;
    return apps::fetch_output(in->c, in->conf.matrixDim * in->conf.matrixDim * sizeof(float), checksum);
}

extern "C" int ua_mm_get_inputs(void* h, float* A, float* B) {
    Instance* in = (Instance*)h;
    if (!in || !A || !B) return -1;
    const size_t bytes = in->conf.matrixDim * in->conf.matrixDim * sizeof(float);
    pmMemHandle m[2] = {in->a, in->b};
    float* dst[2] = {A, B};
    for (int i = 0; i < 2; ++i) {
        pmStatus s = pmFetchMemory(m[i]);
        pmRawMemPtr p;
        if (s == pmSuccess) s = pmGetRawMemPtr(m[i], &p);
        if (s != pmSuccess) return (int)s;
        memcpy(dst[i], p, bytes);
    }
    return 0;
}

extern "C" int ua_mm_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    pmReleaseMemory(in->c);
    delete in;    // A and B go with the last holder of `inputs`
    return 0;
}
