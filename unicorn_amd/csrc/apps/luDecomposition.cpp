// luDecomposition app on the pm:: API (reference testSuite/luDecomposition/source/code/src/
// luDecomposition.cpp and luDecompositionCUDA.cu), plus the solve with FP64-residual iterative
// refinement that the benchmark configuration adds.
//
// Factorisation (reference luDecomposition.cpp:385-493): column-major N x N matrix in blocks of
// B x B; for each block step k a task per stage:
//   S1 (1 subtask):              LU of the diagonal block (k,k) (L with the pivots, U unit upper);
//   S2 (2 (nb-k-1) subtasks):    first half U01 = L00^-1 A01 on block (k, k+1+i), second half
//                                L10 = A10 U00^-1 on block (k+1+i, k);
//   S3 ((nb-k-1)^2 subtasks):    A11 -= L10 U01 on block (k+1+r, k+1+c);
// then S1 on the last diagonal block. Subscriptions and the NATURAL/COMPACT offsets follow the
// reference exactly (luDecomposition.cpp:127-347). GPU callbacks call the gfx950 kernels in place
// of cuBLAS: uh_lu_panel, uh_trsm_lln / uh_trsm_run, uh_gemm (split-BF16 or fp32).
//
// Solve (new): Ax = b for nrhs <= 8 right-hand sides. A pristine copy of A is kept in its own
// address space. Each refinement task computes r = b - A x in FP64, solves L U z = r in fp32
// with the factors, and updates x += z in FP64; the host stops when the normwise backward error
// max|r| / (||A||inf max|x| + max|b|) reaches fp32 unit roundoff or after `max_steps`.
#include "app_common.h"
#include "unicorn_hip.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace luDecomposition {

using namespace pm;

enum stage { LU_DECOMPOSE = 0, HORZ_VERT_COMP = 1, DIAGONAL_COMP = 2, REFINE_INIT = 3, REFINE_STEP = 4 };

struct luTaskConf {
    size_t matrixDim;
    size_t blockDim;
    size_t taskStep;      // block step k (the reference passes it as pmTaskInfo.taskId)
    int precision;        // trailing update: 0 split-BF16 MFMA, 2 fp32
    int nrhs;
    int part;             // stage 3: 0 = every trailing block, 1 = block (k+1, k+1) only,
                          // 2 = every trailing block but (k+1, k+1)
};

// Stage-3 subtask id -> (row, col) of its block in the trailing matrix (m x m blocks).
static void s3_block_of(const luTaskConf* c, ulong id, size_t* row, size_t* col) {
    const size_t m = c->matrixDim / c->blockDim - c->taskStep - 1;
    const size_t e = c->part == 2 ? id + 1 : id;
    *row = e / m;
    *col = e % m;
}

#define BLOCK_OFFSET_IN_ELEMS(c, blockRow, blockCol) (((blockRow) + (blockCol) * (c)->matrixDim) * (c)->blockDim)

// one B x B block (column-major, B records of B floats, step N floats)
static void subscribe_block(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si, const luTaskConf* c, size_t br,
                            size_t bc, pmSubscriptionType type) {
    const size_t off = BLOCK_OFFSET_IN_ELEMS(c, br, bc) * sizeof(float);
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 0, type,
                        pmScatteredSubscriptionInfo(off, c->blockDim * sizeof(float), c->matrixDim * sizeof(float),
                                                    c->blockDim));
}

static pmStatus s1DataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    subscribe_block(ti, di, si, c, c->taskStep, c->taskStep, READ_WRITE_SUBSCRIPTION);
    return pmSuccess;
}

static pmStatus s2DataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    const size_t k = c->taskStep, half = ti.subtaskCount / 2;
    subscribe_block(ti, di, si, c, k, k, READ_SUBSCRIPTION);
    if (si.subtaskId < half) subscribe_block(ti, di, si, c, k, k + 1 + si.subtaskId, READ_WRITE_SUBSCRIPTION);
    else subscribe_block(ti, di, si, c, k + 1 + si.subtaskId - half, k, READ_WRITE_SUBSCRIPTION);
    return pmSuccess;
}

static pmStatus s3DataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    const size_t k = c->taskStep;
    size_t row, col;
    s3_block_of(c, si.subtaskId, &row, &col);
    subscribe_block(ti, di, si, c, k + 1 + row, k, READ_SUBSCRIPTION);
    subscribe_block(ti, di, si, c, k, k + 1 + col, READ_SUBSCRIPTION);
    subscribe_block(ti, di, si, c, k + 1 + row, k + 1 + col, READ_WRITE_SUBSCRIPTION);
    return pmSuccess;
}

static bool natural(const pmSubtaskInfo& si) { return si.memInfo[0].visibilityType == SUBSCRIPTION_NATURAL; }

// ---- CPU callbacks (float loops in the reference's operation order) ----

static void lu_block_cpu(float* a, size_t n, size_t ld) {
    for (size_t i = 0; i + 1 < n; ++i) {
        const float d = (float)(1.0 / (double)a[i + i * ld]);
        for (size_t c = i + 1; c < n; ++c) a[i + c * ld] *= d;
        for (size_t c = i + 1; c < n; ++c) {
            const float u = a[i + c * ld];
            for (size_t r = i + 1; r < n; ++r) a[r + c * ld] -= a[r + i * ld] * u;
        }
    }
}

static pmStatus s1_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    lu_block_cpu((float*)si.memInfo[0].ptr, c->blockDim, natural(si) ? c->matrixDim : c->blockDim);
    return pmSuccess;
}

struct S2Views { float* tri; float* x; size_t ld; bool upper; };

static S2Views s2_views(const luTaskConf* c, pmTaskInfo ti, const pmSubtaskInfo& si) {
    const size_t k = c->taskStep, half = ti.subtaskCount / 2, B = c->blockDim;
    S2Views v;
    v.upper = si.subtaskId < half;
    float* base = (float*)si.memInfo[0].ptr;
    size_t off;
    if (v.upper) {
        // luDecomposition.cpp:237-251
        if (natural(si)) { off = BLOCK_OFFSET_IN_ELEMS(c, k, k + 1 + si.subtaskId) - BLOCK_OFFSET_IN_ELEMS(c, k, k); v.ld = c->matrixDim; }
        else { off = B * B; v.ld = B; }
    } else {
        // luDecomposition.cpp:257-271
        if (natural(si)) { off = BLOCK_OFFSET_IN_ELEMS(c, k + 1 + si.subtaskId - half, k) - BLOCK_OFFSET_IN_ELEMS(c, k, k); v.ld = c->matrixDim; }
        else { off = B; v.ld = 2 * B; }
    }
    v.tri = base;
    v.x = base + off;
    return v;
}

static pmStatus s2_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    S2Views v = s2_views(c, ti, si);
    const size_t B = c->blockDim, ld = v.ld;
    if (v.upper) {          // X := L^-1 X, L lower non-unit
        for (size_t col = 0; col < B; ++col)
            for (size_t i = 0; i < B; ++i) {
                float s = v.x[i + col * ld];
                for (size_t j = 0; j < i; ++j) s -= v.tri[i + j * ld] * v.x[j + col * ld];
                v.x[i + col * ld] = s / v.tri[i + i * ld];
            }
    } else {                // X := X U^-1, U upper unit
        for (size_t row = 0; row < B; ++row)
            for (size_t j = 0; j < B; ++j) {
                float s = v.x[row + j * ld];
                for (size_t i = 0; i < j; ++i) s -= v.x[row + i * ld] * v.tri[i + j * ld];
                v.x[row + j * ld] = s;
            }
    }
    return pmSuccess;
}

struct S3Views { float* l10; float* u01; float* a11; size_t ld1, ld2, ld3; };

static S3Views s3_views(const luTaskConf* c, pmTaskInfo ti, const pmSubtaskInfo& si) {
    const size_t k = c->taskStep, B = c->blockDim;
    size_t row, col;
    s3_block_of(c, si.subtaskId, &row, &col);
    S3Views v;
    float* base = (float*)si.memInfo[0].ptr;
    size_t o1, o2;
    if (natural(si)) {   // luDecomposition.cpp:322-329
        o1 = BLOCK_OFFSET_IN_ELEMS(c, k, k + 1 + col) - BLOCK_OFFSET_IN_ELEMS(c, k + 1 + row, k);
        o2 = BLOCK_OFFSET_IN_ELEMS(c, k + 1 + row, k + 1 + col) - BLOCK_OFFSET_IN_ELEMS(c, k + 1 + row, k);
        v.ld1 = v.ld2 = v.ld3 = c->matrixDim;
    } else {             // luDecomposition.cpp:330-337
        o1 = B * B;
        o2 = B * B + B;
        v.ld1 = B;
        v.ld2 = v.ld3 = 2 * B;
    }
    v.l10 = base;
    v.u01 = base + o1;
    v.a11 = base + o2;
    return v;
}

static pmStatus s3_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    S3Views v = s3_views(c, ti, si);
    const size_t B = c->blockDim;
    for (size_t j = 0; j < B; ++j)
        for (size_t kk = 0; kk < B; ++kk) {
            const float u = v.u01[kk + j * v.ld2];
            for (size_t i = 0; i < B; ++i) v.a11[i + j * v.ld3] -= v.l10[i + kk * v.ld1] * u;
        }
    return pmSuccess;
}

// ---- GPU callbacks ----

static pmStatus gpu_error(const char* what) {
    apps::set_error(std::string(what) + ": " + uh_last_error());
    return pmGraphicsCardError;
}

static pmStatus s1_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    if (uh_lu_panel((float*)si.memInfo[0].ptr, natural(si) ? c->matrixDim : c->blockDim, (int)c->blockDim, stream))
        return gpu_error("uh_lu_panel");
    return pmSuccess;
}

static pmStatus s2_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    S2Views v = s2_views(c, ti, si);
    const int B = (int)c->blockDim;
    int rc = v.upper ? uh_trsm_lln(v.tri, v.ld, v.x, v.ld, B, B, stream)
                     : uh_trsm_run(v.tri, v.ld, v.x, v.ld, B, B, stream);
    if (rc) return gpu_error("uh_trsm");
    return pmSuccess;
}

static pmStatus s3_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    S3Views v = s3_views(c, ti, si);
    const int B = (int)c->blockDim;
    // column-major A11 -= L10 U01  <=>  row-major A11^T -= U01^T L10^T
    uh_gemm_args g;
    g.A = v.u01;
    g.lda = v.ld2;
    g.B = v.l10;
    g.ldb = v.ld1;
    g.C = v.a11;
    g.ldc = v.ld3;
    g.M = B;
    g.N = B;
    g.K = B;
    g.precision = c->precision;
    g.alpha = -1.f;
    g.beta = 1.f;
    if (uh_gemm(&g, stream)) return gpu_error("uh_gemm");
    return pmSuccess;
}

// ---- solve with iterative refinement ----
// Address spaces: 0 = LU factors (READ_ONLY), 1 = pristine A (READ_ONLY), 2 = solve state
// (READ_WRITE): [b | x | r] n x nrhs float64 column-major each, [rf] n x nrhs float32, then
// [norms] 2 * 8 float64 (max|r| per column, max|x| per column). Between steps the host reads
// only the norms (pmReadMemoryRange), so the state stays resident in HBM until x is fetched.

struct SolveLayout {
    size_t n, nrhs;
    size_t b() const { return 0; }
    size_t x() const { return n * nrhs * 8; }
    size_t r() const { return 2 * n * nrhs * 8; }
    size_t rf() const { return 3 * n * nrhs * 8; }
    size_t norms() const { return 3 * n * nrhs * 8 + n * nrhs * 4; }
    size_t bytes() const { return norms() + 16 * 8; }
};

static pmStatus solveDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    const size_t N = c->matrixDim;
    SolveLayout L{N, (size_t)c->nrhs};
    if (c->taskStep == REFINE_STEP) {
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 0, READ_SUBSCRIPTION,
                            pmSubscriptionInfo(0, N * N * 4));
        pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 1, READ_SUBSCRIPTION,
                            pmSubscriptionInfo(0, N * N * 4));
    }
    // REFINE_INIT: x = 0, r = b, rf = (float) b. REFINE_STEP: solve L U z = rf, x += z, then
    // r = b - A x (also as rf, the next step's right-hand side) and the norms.
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, 2, READ_WRITE_SUBSCRIPTION,
                        pmSubscriptionInfo(0, L.bytes()));
    return pmSuccess;
}

static pmStatus solve_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    const int N = (int)c->matrixDim, k = c->nrhs;
    SolveLayout L{(size_t)N, (size_t)k};
    uint8_t* st = (uint8_t*)si.memInfo[2].ptr;
    double* b = (double*)(st + L.b());
    double* x = (double*)(st + L.x());
    double* r = (double*)(st + L.r());
    float* rf = (float*)(st + L.rf());
    if (c->taskStep == REFINE_INIT) {
        if (uh_refine_init(b, x, r, rf, N, k, stream)) return gpu_error("uh_refine_init");
        return pmSuccess;
    }
    const float* lu = (const float*)si.memInfo[0].ptr;
    const float* a = (const float*)si.memInfo[1].ptr;
    double* norms = (double*)(st + L.norms());
    if (uh_lu_solve_f32(lu, (uint64_t)N, N, rf, k, stream)) return gpu_error("uh_lu_solve_f32");
    if (uh_refine_update(x, rf, N, k, stream)) return gpu_error("uh_refine_update");
    // residual of the updated x (the next step's right-hand side), and the norms the host uses
    // to decide convergence
    if (uh_residual_f64(a, (uint64_t)N, N, x, b, k, r, rf, stream)) return gpu_error("uh_residual_f64");
    if (uh_colmax_abs_f64(r, N, k, norms, stream)) return gpu_error("norms");
    if (uh_colmax_abs_f64(x, N, k, norms + 8, stream)) return gpu_error("norms");
    return pmSuccess;
}

static pmStatus solve_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const luTaskConf* c = (const luTaskConf*)ti.taskConf;
    const size_t N = c->matrixDim, k = (size_t)c->nrhs;
    SolveLayout L{N, k};
    uint8_t* st = (uint8_t*)si.memInfo[2].ptr;
    double* b = (double*)(st + L.b());
    double* x = (double*)(st + L.x());
    double* r = (double*)(st + L.r());
    float* rf = (float*)(st + L.rf());
    if (c->taskStep == REFINE_INIT) {
        memset(x, 0, N * k * 8);
        for (size_t i = 0; i < N * k; ++i) {
            r[i] = b[i];
            rf[i] = (float)b[i];
        }
        return pmSuccess;
    }
    const float* lu = (const float*)si.memInfo[0].ptr;
    const float* a = (const float*)si.memInfo[1].ptr;
    double* norms = (double*)(st + L.norms());
    auto residual = [&]() {
        for (size_t q = 0; q < k; ++q) {
            for (size_t i = 0; i < N; ++i) r[q * N + i] = b[q * N + i];
            for (size_t j = 0; j < N; ++j) {
                const double xj = x[q * N + j];
                for (size_t i = 0; i < N; ++i) r[q * N + i] -= (double)a[i + j * N] * xj;
            }
            for (size_t i = 0; i < N; ++i) rf[q * N + i] = (float)r[q * N + i];
        }
    };
    for (size_t q = 0; q < k; ++q) {
        float* z = rf + q * N;      // r of the previous step (or of x = 0), in fp32
        for (size_t i = 0; i < N; ++i) {
            float s = z[i];
            for (size_t j = 0; j < i; ++j) s -= lu[i + j * N] * z[j];
            z[i] = s / lu[i + i * N];
        }
        for (size_t i = N; i-- > 0;) {
            float s = z[i];
            for (size_t j = i + 1; j < N; ++j) s -= lu[i + j * N] * z[j];
            z[i] = s;
        }
        for (size_t i = 0; i < N; ++i) x[q * N + i] += (double)z[i];
    }
    residual();
    for (size_t q = 0; q < k; ++q) {
        double mr = 0, mx = 0;
        for (size_t i = 0; i < N; ++i) {
            mr = std::max(mr, std::fabs(r[q * N + i]));
            mx = std::max(mx, std::fabs(x[q * N + i]));
        }
        norms[q] = mr;
        norms[8 + q] = mx;
    }
    return pmSuccess;
}

struct Instance {
    size_t n = 0, block = 0;
    int precision = 0;
    pmMemHandle mat = nullptr, orig = nullptr, state = nullptr;
    double a_inf = 0;    // ||A||inf of the pristine matrix
};

static apps::CallbackVariants g_cb[4];
static std::mutex g_mtx;

static int register_all() {
    std::lock_guard<std::mutex> g(g_mtx);
    pmCallbacks c1, c2, c3, c4;
    c1.dataDistribution = s1DataDistribution; c1.subtask_cpu = s1_cpu; c1.subtask_gpu_custom = s1_gpu;
    c2.dataDistribution = s2DataDistribution; c2.subtask_cpu = s2_cpu; c2.subtask_gpu_custom = s2_gpu;
    c3.dataDistribution = s3DataDistribution; c3.subtask_cpu = s3_cpu; c3.subtask_gpu_custom = s3_gpu;
    c4.dataDistribution = solveDataDistribution; c4.subtask_cpu = solve_cpu; c4.subtask_gpu_custom = solve_gpu;
    if (apps::register_variants("LU_S1", c1, g_cb[0]) || apps::register_variants("LU_S2", c2, g_cb[1]) ||
        apps::register_variants("LU_S3", c3, g_cb[2]) || apps::register_variants("LU_SOLVE", c4, g_cb[3])) {
        apps::set_error(std::string("register callbacks: ") + pmGetLastError());
        return -1;
    }
    return 0;
}

// One stage task, submitted (not waited for): conf must outlive it.
static int submit_stage(Instance* in, int stage, size_t k, int part, int device, luTaskConf* conf,
                        pmTaskHandle* th, const pmTaskSubmitOptions* opt = nullptr) {
    conf->matrixDim = in->n;
    conf->blockDim = in->block;
    conf->taskStep = k;
    conf->precision = in->precision;
    conf->nrhs = 0;
    conf->part = part;
    const size_t nb = in->n / in->block, m = nb - k - 1;
    const ulong subtasks = stage == LU_DECOMPOSE ? 1
                         : stage == HORZ_VERT_COMP ? 2 * m
                         : part == 1 ? 1 : part == 2 ? m * m - 1 : m * m;
    pmTaskMem mem[1];
    // reference luDecomposition.cpp:418: READ_WRITE, OPTIMAL visibility, disjoint read/writes
    mem[0] = pmTaskMem(in->mat, READ_WRITE, SUBSCRIPTION_OPTIMAL, true);
    pmTaskDetails d(conf, sizeof(*conf), mem, 1, g_cb[stage].h[device], subtasks);
    d.taskId = k;
    return (int)pmSubmitTaskWithOptions(d, opt, th);
}

static int wait_stage(pmTaskHandle th) {
    const int rc = (int)pmWaitForTaskCompletion(th);
    pmReleaseTask(th);
    return rc;
}

static int stage_task(Instance* in, int stage, size_t k, int device, double* secs, int part = 0) {
    luTaskConf conf;
    pmTaskHandle th = nullptr;
    const double t0 = apps::now_secs();
    int rc = submit_stage(in, stage, k, part, device, &conf, &th);
    if (!rc) rc = wait_stage(th);
    if (secs) *secs = apps::now_secs() - t0;
    return rc;
}

}  // namespace luDecomposition

using namespace luDecomposition;

extern "C" void* ua_lu_create(int n, int block, int precision) {
    if (n <= 0 || block <= 0 || n % block || block > 512 || (precision != 0 && precision != 2)) {
        apps::set_error("ua_lu_create: n must be a multiple of block <= 512; precision 0 or 2");
        return nullptr;
    }
    if (register_all()) return nullptr;
    Instance* in = new Instance();
    in->n = (size_t)n;
    in->block = (size_t)block;
    in->precision = precision;
    const size_t bytes = in->n * in->n * 4;
    if (pmCreateMemory2D(in->n, in->n * 4, &in->mat) != pmSuccess || pmCreateMemory2D(in->n, in->n * 4, &in->orig) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    (void)bytes;
    // Stage 3 runs as batches of (nb-k-1)^2 independent B^3 products, so the tile is tuned on
    // one batch seen as a single product (g B square, K = B, g = min(nb - 1, 8)) and recorded for
    // the B^3 shape; with one trailing block the single product is the shape.
    const int g = std::min(n / block - 1, 8);
    if (g >= 2) apps::autotune_gemm_as(g * block, g * block, block, block, block, block, precision);
    else apps::autotune_gemm(block, block, block, precision);
    return in;
}

// A: n x n float32, column-major.
extern "C" int ua_lu_set_matrix(void* h, const float* A) {
    Instance* in = (Instance*)h;
    const size_t bytes = in->n * in->n * 4;
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->mat, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, A, bytes);
    s = pmGetRawMemPtr(in->orig, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, A, bytes);
    double m = 0;
    for (size_t i = 0; i < in->n; ++i) {
        double rs = 0;
        for (size_t j = 0; j < in->n; ++j) rs += std::fabs((double)A[i + j * in->n]);
        m = std::max(m, rs);
    }
    in->a_inf = m;
    return 0;
}

// Makes the matrix and its pristine copy resident in HBM ahead of the factorisation.
extern "C" int ua_lu_prefetch(void* h) {
    Instance* in = (Instance*)h;
    if (!pmHasHipDevice()) return 0;
    pmStatus s = pmPrefetchToDevice(in->mat);
    if (s == pmSuccess) s = pmPrefetchToDevice(in->orig);
    return (int)s;
}

// Blocked factorisation: 3 tasks per block step (reference luDecomposition.cpp:466-480).
//
// On the GPU in the batched mode the driver looks one step ahead: stage 3 of step k first
// updates block (k+1, k+1) alone, then the factorisation of that block (stage 1 of step k+1, a
// chain of small launches, submitted as a RESIDENT-mode task so that its launches go to a stream
// of their own) and the rest of step k's trailing update (batch stream) run together: the stub
// stages the update while the chain runs. The two tasks write disjoint blocks and read none of
// each other's, so the update is submitted unordered (it does not wait for the factorisation,
// which shares its address space); both carry their stub mode in their submit options.
// (Measured: 47.6 ms against 49.4 ms at N = 16384; the GPU overlaps little, because the
// persistent update holds every CU while it runs.)
extern "C" int ua_lu_factor(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    const size_t nb = in->n / in->block;
    double t = 0, s = 0;
    if (device == 1 && pmGetHipStubMode() == PM_HIP_MODE_BATCHED && nb >= 2) {
        const double t0 = apps::now_secs();
        int rc = stage_task(in, LU_DECOMPOSE, 0, device, nullptr);
        for (size_t k = 0; rc == 0 && k + 1 < nb; ++k) {
            rc = stage_task(in, HORZ_VERT_COMP, k, device, nullptr);
            if (!rc) rc = stage_task(in, DIAGONAL_COMP, k, device, nullptr, 1);
            if (rc) break;
            const size_t m = nb - k - 1;
            luTaskConf c1, c3;
            pmTaskHandle h1 = nullptr, h3 = nullptr;
            const pmTaskSubmitOptions chain{PM_HIP_MODE_RESIDENT, false};
            const pmTaskSubmitOptions update{PM_HIP_MODE_BATCHED, true};
            rc = submit_stage(in, LU_DECOMPOSE, k + 1, 0, device, &c1, &h1, &chain);
            if (!rc && m >= 2) rc = submit_stage(in, DIAGONAL_COMP, k, 2, device, &c3, &h3, &update);
            const int r1 = h1 ? wait_stage(h1) : 0;
            const int r3 = h3 ? wait_stage(h3) : 0;
            if (!rc) rc = r1 ? r1 : r3;
        }
        if (seconds) *seconds = apps::now_secs() - t0;
        return rc;
    }
    for (size_t k = 0; k + 1 < nb; ++k) {
        int rc = stage_task(in, LU_DECOMPOSE, k, device, &s);
        t += s;
        if (!rc) { rc = stage_task(in, HORZ_VERT_COMP, k, device, &s); t += s; }
        if (!rc) { rc = stage_task(in, DIAGONAL_COMP, k, device, &s); t += s; }
        if (rc) return rc;
    }
    int rc = stage_task(in, LU_DECOMPOSE, nb - 1, device, &s);
    t += s;
    if (seconds) *seconds = t;
    return rc;
}

extern "C" int ua_lu_get_factors(void* h, float* LU) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->mat);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->mat, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(LU, p, in->n * in->n * 4);
    return 0;
}

// Solve with the current factors. b, x: n x nrhs float64 column-major. info[0] = refinement
// steps run, info[1] = final normwise backward error (max over right-hand sides), info[2] =
// backward error after the first step, info[3] = ||A||inf.
extern "C" int ua_lu_solve(void* h, const double* b, int nrhs, int max_steps, int device, double* x, double* info,
                           double* seconds) {
    Instance* in = (Instance*)h;
    if (nrhs < 1 || nrhs > 8 || max_steps < 1 || device < 0 || device > 2) return -1;
    const size_t N = in->n;
    SolveLayout L{N, (size_t)nrhs};
    if (!in->state || true) {
        if (in->state) pmReleaseMemory(in->state);
        in->state = nullptr;
        if (pmCreateMemory(L.bytes(), &in->state) != pmSuccess) return (int)pmMemoryError;
    }
    pmRawMemPtr p;
    pmStatus st = pmGetRawMemPtr(in->state, &p);
    if (st != pmSuccess) return (int)st;
    memset(p, 0, L.bytes());
    memcpy((uint8_t*)p + L.b(), b, N * nrhs * 8);
    double bmax = 0;
    for (size_t i = 0; i < N * (size_t)nrhs; ++i) bmax = std::max(bmax, std::fabs(b[i]));
    luTaskConf conf;
    conf.matrixDim = N;
    conf.blockDim = in->block;
    conf.precision = in->precision;
    conf.nrhs = nrhs;
    pmTaskMem mem[3];
    mem[0] = pmTaskMem(in->mat, READ_ONLY, SUBSCRIPTION_NATURAL);
    mem[1] = pmTaskMem(in->orig, READ_ONLY, SUBSCRIPTION_NATURAL);
    mem[2] = pmTaskMem(in->state, READ_WRITE, SUBSCRIPTION_NATURAL, true);   // one subtask: disjoint
    double t = 0, s = 0;
    conf.taskStep = REFINE_INIT;
    pmTaskDetails d(&conf, sizeof(conf), mem, 3, g_cb[3].h[device], 1);
    int rc = apps::run_task(d, &s);
    t += s;
    const double eps = std::ldexp(1.0, -24);
    int steps = 0;
    double first = -1, berr = 1;
    for (; rc == 0 && steps < max_steps;) {
        conf.taskStep = REFINE_STEP;
        pmTaskDetails ds(&conf, sizeof(conf), mem, 3, g_cb[3].h[device], 1);
        rc = apps::run_task(ds, &s);
        t += s;
        ++steps;
        if (rc) break;
        const double t0 = apps::now_secs();
        double norms[16];
        st = pmReadMemoryRange(in->state, L.norms(), sizeof(norms), norms);
        if (st != pmSuccess) { rc = (int)st; break; }
        berr = 0;
        for (int q = 0; q < nrhs; ++q) berr = std::max(berr, norms[q] / (in->a_inf * norms[8 + q] + bmax));
        t += apps::now_secs() - t0;
        if (first < 0) first = berr;
        if (berr <= eps) break;
    }
    if (rc) return rc;
    st = pmFetchMemory(in->state);
    if (st != pmSuccess) return (int)st;
    pmGetRawMemPtr(in->state, &p);
    memcpy(x, (uint8_t*)p + L.x(), N * nrhs * 8);
    if (info) { info[0] = steps; info[1] = berr; info[2] = first; info[3] = in->a_inf; }
    if (seconds) *seconds = t;
    return 0;
}

extern "C" int ua_lu_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    pmReleaseMemory(in->mat);
    pmReleaseMemory(in->orig);
    if (in->state) pmReleaseMemory(in->state);
    delete in;
    return 0;
}
