// fft app (reference testSuite/fft as built: FFT_2D with NO_MATRIX_TRANSPOSE and USE_SQUARE_MATRIX,
// fft.h:33-60): the forward, unnormalised 2-D DFT of a rows x cols matrix of complex floats,
//   X[u][v] = sum_{r,c} x[r][c] exp(-2 pi i (r u / rows + c v / cols)),
// as two tasks over two address spaces (DoParallelProcess / Parallel_FFT_1D, fft.cpp:401-563):
//   row task: subtask s transforms rows [s R, s R + R) (R = ROWS_PER_FFT_SUBTASK, 512 by default),
//     input READ_ONLY -> output WRITE_ONLY, contiguous NATURAL subscriptions;
//   column task: subtask s transforms columns [s R, s R + R) of the row task's output and writes
//     them into the first space: scattered subscriptions of R columns x rows, OPTIMAL visibility;
//   the result is left in the first space (the two handles swap, fft.cpp:529-534).
// fftDataDistribution (fft.cpp:157-284): CPU subtasks can be split (canSplitCpuSubtasks): split
// k of K takes R / K rows (columns), the last split the rest. The reference subscribes elemsX / R
// column subtasks, which covers every column only for its square matrices; here the column task
// has cols / R subtasks, so rows != cols works too (both powers of two, R dividing both).
// CPU callback fft_cpu (fft.cpp:286-322) ran FFTW plans; here an iterative radix-2 transform in
// float with twiddles rounded once from double. GPU callback fft_cudaLaunchFunc
// (fftCUDA.cu:182-205) ran cuFFT plans; here uh_fft_c2c_f32 over the subtask's rows, or over its
// columns with the row stride of the view it got (NATURAL: the matrix row, COMPACT: R columns).
// In-place 2-D transforms are not supported by the reference under NO_MATRIX_TRANSPOSE
// (fft.cpp:599-603), nor here.
#include "app_common.h"

#include <cmath>
#include <complex>

namespace fft {

using namespace pm;

enum memIndex { INPUT_MEM_INDEX = 0, OUTPUT_MEM_INDEX, MAX_MEM_INDICES };
static constexpr size_t ROWS_PER_FFT_SUBTASK = 512;

// reference fftTaskConf (fft.h:86-92) plus the subtask width, which the reference fixes at 512
struct fftTaskConf {
    size_t elemsX;    // rows
    size_t elemsY;    // cols
    bool rowPlanner;
    bool inplace;
    size_t rowsPerSubtask;
};

typedef std::complex<float> cfloat;

// Rows (columns) of split k of K of a subtask of R: R / K each, the last one the rest.
static void split_range(const fftTaskConf* c, const pmSplitInfo& sp, size_t& first, size_t& count) {
    const size_t R = c->rowsPerSubtask;
    if (!sp.splitCount) {
        first = 0;
        count = R;
        return;
    }
    const size_t per = R / sp.splitCount;
    first = per * sp.splitId;
    count = sp.splitId == sp.splitCount - 1 ? R - first : per;
}

static pmStatus fftDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const fftTaskConf* c = (const fftTaskConf*)ti.taskConf;
    size_t first, count;
    if (si.splitInfo.splitCount > c->rowsPerSubtask) return pmInvalidIndex;
    split_range(c, si.splitInfo, first, count);
    const size_t start = c->rowsPerSubtask * si.subtaskId + first;
    pmStatus st;
    if (c->rowPlanner) {
        const pmSubscriptionInfo s(start * c->elemsY * sizeof(cfloat), count * c->elemsY * sizeof(cfloat));
        st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MEM_INDEX,
                                 READ_SUBSCRIPTION, s);
        if (st == pmSuccess)
            st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_MEM_INDEX,
                                     WRITE_SUBSCRIPTION, s);
    } else {
        const pmScatteredSubscriptionInfo s(start * sizeof(cfloat), count * sizeof(cfloat), c->elemsY * sizeof(cfloat),
                                            c->elemsX);
        st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MEM_INDEX,
                                 READ_SUBSCRIPTION, s);
        if (st == pmSuccess)
            st = pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_MEM_INDEX,
                                     WRITE_SUBSCRIPTION, s);
    }
    return st;
}

// ---- CPU transform: iterative radix-2 decimation in time over a contiguous sequence

struct Twiddles {
    size_t n = 0;
    std::vector<cfloat> w;     // exp(-2 pi i k / n), k < n / 2
    std::vector<uint32_t> rev; // bit reversal permutation
};

static const Twiddles& twiddles(size_t n) {
    static std::mutex m;
    static std::vector<std::unique_ptr<Twiddles>> cache;
    std::lock_guard<std::mutex> g(m);
    for (auto& t : cache)
        if (t->n == n) return *t;
    std::unique_ptr<Twiddles> t(new Twiddles());
    t->n = n;
    t->w.resize(n / 2);
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t k = 0; k < n / 2; ++k) {
        const double a = two_pi * (double)k / (double)n;
        t->w[k] = cfloat((float)std::cos(a), (float)-std::sin(a));
    }
    t->rev.resize(n);
    int bits = 0;
    while (((size_t)1 << bits) < n) ++bits;
    for (size_t i = 0; i < n; ++i) {
        uint32_t r = 0;
        for (int b = 0; b < bits; ++b)
            if (i >> b & 1) r |= 1u << (bits - 1 - b);
        t->rev[i] = r;
    }
    cache.push_back(std::move(t));
    return *cache.back();
}

// y = DFT(x) of n points; x is read with stride xs, the result written to y (contiguous, n).
static void fft_seq(const cfloat* x, size_t xs, cfloat* y, const Twiddles& tw) {
    const size_t n = tw.n;
    for (size_t i = 0; i < n; ++i) y[tw.rev[i]] = x[i * xs];
    for (size_t half = 1; half < n; half <<= 1) {
        const size_t step = n / (2 * half);
        for (size_t base = 0; base < n; base += 2 * half)
            for (size_t k = 0; k < half; ++k) {
                const cfloat w = tw.w[k * step];
                const cfloat a = y[base + k];
                const cfloat b = y[base + k + half];
                const cfloat t(w.real() * b.real() - w.imag() * b.imag(), w.real() * b.imag() + w.imag() * b.real());
                y[base + k] = a + t;
                y[base + k + half] = a - t;
            }
    }
}

static pmStatus fft_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const fftTaskConf* c = (const fftTaskConf*)ti.taskConf;
    size_t first, count;
    split_range(c, si.splitInfo, first, count);
    const cfloat* in = (const cfloat*)si.memInfo[INPUT_MEM_INDEX].ptr;
    cfloat* out = (cfloat*)si.memInfo[OUTPUT_MEM_INDEX].ptr;
    if (c->rowPlanner) {
        const Twiddles& tw = twiddles(c->elemsY);
        for (size_t r = 0; r < count; ++r) fft_seq(in + r * c->elemsY, 1, out + r * c->elemsY, tw);
    } else {
        // NATURAL views: column j of the subtask at [j], row stride elemsY; a COMPACT view packs
        // the subscribed columns of a row back to back
        const Twiddles& tw = twiddles(c->elemsX);
        const size_t ldi = si.memInfo[INPUT_MEM_INDEX].visibilityType == SUBSCRIPTION_COMPACT ? count : c->elemsY;
        const size_t ldo = si.memInfo[OUTPUT_MEM_INDEX].visibilityType == SUBSCRIPTION_COMPACT ? count : c->elemsY;
        std::vector<cfloat> col(c->elemsX);
        for (size_t j = 0; j < count; ++j) {
            fft_seq(in + j, ldi, col.data(), tw);
            for (size_t r = 0; r < c->elemsX; ++r) out[j + r * ldo] = col[r];
        }
    }
    return pmSuccess;
}

static pmStatus fft_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const fftTaskConf* c = (const fftTaskConf*)ti.taskConf;
    size_t first, count;
    split_range(c, si.splitInfo, first, count);
    const void* in = si.memInfo[INPUT_MEM_INDEX].ptr;
    void* out = si.memInfo[OUTPUT_MEM_INDEX].ptr;
    int rc;
    if (c->rowPlanner) {
        rc = uh_fft_c2c_f32(in, c->elemsY, 1, out, c->elemsY, 1, (uint32_t)c->elemsY, (uint32_t)count, stream);
    } else {
        const size_t ldi = si.memInfo[INPUT_MEM_INDEX].visibilityType == SUBSCRIPTION_COMPACT ? count : c->elemsY;
        const size_t ldo = si.memInfo[OUTPUT_MEM_INDEX].visibilityType == SUBSCRIPTION_COMPACT ? count : c->elemsY;
        rc = uh_fft_c2c_f32(in, 1, ldi, out, 1, ldo, (uint32_t)c->elemsX, (uint32_t)count, stream);
    }
    if (rc != 0) {
        apps::set_error(std::string("uh_fft_c2c_f32: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    fftTaskConf conf;
    pmMemHandle a = nullptr, b = nullptr;   // a: input, then the result; b: the row task's output
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

static bool pow2(size_t v) { return v && !(v & (v - 1)); }

}  // namespace fft

using namespace fft;

extern "C" void* ua_fft_create(uint32_t rows, uint32_t cols, uint32_t rows_per_subtask) {
    const size_t R = rows_per_subtask ? rows_per_subtask : ROWS_PER_FFT_SUBTASK;
    if (!pow2(rows) || !pow2(cols) || rows < 2 || cols < 2 || rows > 8192 || cols > 8192) {
        apps::set_error("ua_fft_create: rows and cols must be powers of two in [2, 8192]");
        return nullptr;
    }
    if (!pow2(R) || R > rows || R > cols) {
        apps::set_error("ua_fft_create: rows per subtask must be a power of two dividing rows and cols");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = fftDataDistribution;
        cb.subtask_cpu = fft_cpu;
        cb.subtask_gpu_custom = fft_gpu;
        if (apps::register_variants("FFT", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    in->conf.elemsX = rows;
    in->conf.elemsY = cols;
    in->conf.rowPlanner = true;
    in->conf.inplace = false;
    in->conf.rowsPerSubtask = R;
    if (pmCreateMemory2D(rows, (size_t)cols * sizeof(cfloat), &in->a) != pmSuccess ||
        pmCreateMemory2D(rows, (size_t)cols * sizeof(cfloat), &in->b) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        if (in->a) pmReleaseMemory(in->a);
        delete in;
        return nullptr;
    }
    return in;
}

// data: rows x cols complex values as interleaved float pairs
extern "C" int ua_fft_set_input(void* h, const float* data) {
    Instance* in = (Instance*)h;
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->a, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, data, in->conf.elemsX * in->conf.elemsY * sizeof(cfloat));
    return 0;
}

// Row task then column task; seconds = wall time of both.
extern "C" int ua_fft_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    const double t0 = apps::now_secs();
    for (int pass = 0; pass < 2; ++pass) {
        fftTaskConf conf = in->conf;
        conf.rowPlanner = pass == 0;
        const pmMemHandle src = pass == 0 ? in->a : in->b, dst = pass == 0 ? in->b : in->a;
        const pmSubscriptionVisibilityType v = pass == 0 ? SUBSCRIPTION_NATURAL : SUBSCRIPTION_OPTIMAL;
        pmTaskMem mem[MAX_MEM_INDICES] = {pmTaskMem(src, READ_ONLY, v), pmTaskMem(dst, WRITE_ONLY, v)};
        const unsigned long subtasks = (pass == 0 ? conf.elemsX : conf.elemsY) / conf.rowsPerSubtask;
        pmTaskDetails d(&conf, sizeof(conf), mem, MAX_MEM_INDICES, g_cb.h[device], subtasks);
        d.canSplitCpuSubtasks = apps::split_request().cpu;
        d.multiAssignEnabled = apps::split_request().multi_assign;
        const int rc = apps::run_task(d, nullptr);
        if (rc) return rc;
    }
    if (seconds) *seconds = apps::now_secs() - t0;
    return 0;
}

extern "C" int ua_fft_get_output(void* h, float* out) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->a);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->a, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(out, p, in->conf.elemsX * in->conf.elemsY * sizeof(cfloat));
    return 0;
}

extern "C" int ua_fft_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    pmReleaseMemory(in->a);
    pmReleaseMemory(in->b);
    delete in;
    return 0;
}
