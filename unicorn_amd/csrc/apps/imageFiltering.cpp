// imageFiltering app on the pm:: API (reference testSuite/imageFiltering/source/code/src/
// imageFiltering.cpp and imageFilteringCUDA.cu).
//
// Task: one subtask per tile of `tile` x `tile` output pixels (reference TILE_DIM, GetSubtaskSubscription
// imageFiltering.cpp:279-340). Each subtask subscribes to the halo'd input region in the
// address space's inverted row order and to its output tile (imageFilterDataDistribution,
// imageFiltering.cpp:342-377). The CPU callback is the reference's imageFilter_cpu
// (imageFiltering.cpp:379-472) with the second arithmetic mode added; the GPU callback hands the
// subtask's view to the gfx950 kernel (uh_imagefilter). Unlike the reference GPU callback it
// reserves no device memory: the kernel reads the view directly, so there is no texture copy.
#include "app_common.h"
#include "unicorn_hip.h"

#include <vector>

namespace imageFiltering {

using namespace pm;

static constexpr size_t PIXEL_COUNT = 3;
static constexpr int MAX_FILTER_DIM = 32;

enum memIndex { INPUT_MEM_INDEX = 0, OUTPUT_MEM_INDEX = 1, MAX_MEM_INDICES };

// Task configuration: the reference's imageFilterTaskConf (imageFiltering.h:96-109) plus the
// arithmetic mode, the right shift of mode 1 and the tile side.
struct imageFilterTaskConf {
    size_t imageWidth;
    size_t imageHeight;
    size_t imageOffset;
    size_t imageBytesPerLine;
    size_t filterRadius;
    char imagePath[256];
    char filter[MAX_FILTER_DIM][MAX_FILTER_DIM];
    int arithmeticMode;
    int rightShift;
    int tileDim;
};

// Tile of a subtask and its split part (reference imageFiltering.cpp:279-340).
static bool GetSubtaskSubscription(const imageFilterTaskConf* c, unsigned long id, const pmSplitInfo& split,
                                   int* startCol, int* endCol, int* startRow, int* endRow) {
    const unsigned int tilesPerRow = (unsigned int)((c->imageWidth + c->tileDim - 1) / c->tileDim);
    *startCol = (int)((id % tilesPerRow) * c->tileDim);
    *endCol = *startCol + c->tileDim;
    *startRow = (int)((id / tilesPerRow) * c->tileDim);
    *endRow = *startRow + c->tileDim;
    if ((size_t)*endCol > c->imageWidth) *endCol = (int)c->imageWidth;
    if ((size_t)*endRow > c->imageHeight) *endRow = (int)c->imageHeight;
    if (split.splitCount) {
        int count = (int)split.splitCount;
        const int rows = *endRow - *startRow;
        if (count > rows) count = rows;
        if ((int)split.splitId >= count) return false;
        const int factor = rows / count;
        *startRow += (int)split.splitId * factor;
        if ((int)split.splitId != count - 1) *endRow = *startRow + factor;
    }
    return true;
}

struct Halo { int firstRow, lastRow, firstCol, lastCol; };

static Halo halo_of(const imageFilterTaskConf* c, int startCol, int endCol, int startRow, int endRow) {
    const int r = (int)c->filterRadius;
    Halo h;
    h.firstRow = startRow - r < 0 ? 0 : startRow - r;
    h.lastRow = endRow - 1 + r >= (int)c->imageHeight ? (int)c->imageHeight - 1 : endRow - 1 + r;
    h.firstCol = startCol - r < 0 ? 0 : startCol - r;
    h.lastCol = endCol - 1 + r >= (int)c->imageWidth ? (int)c->imageWidth - 1 : endCol - 1 + r;
    return h;
}

static pmStatus imageFilterDataDistribution(pmTaskInfo ti, pmDeviceInfo di, pmSubtaskInfo si) {
    const imageFilterTaskConf* c = (const imageFilterTaskConf*)ti.taskConf;
    int sc, ec, sr, er;
    if (!GetSubtaskSubscription(c, si.subtaskId, si.splitInfo, &sc, &ec, &sr, &er)) return pmSuccess;
    const size_t rowSize = c->imageWidth * PIXEL_COUNT;
    const Halo h = halo_of(c, sc, ec, sr, er);
    // inverted rows: image row m lives at address-space row (H - 1 - m)
    const int firstInv = (int)(c->imageHeight - 1 - h.firstRow);
    const int lastInv = (int)(c->imageHeight - 1 - h.lastRow);
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, INPUT_MEM_INDEX, READ_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo(lastInv * rowSize + h.firstCol * PIXEL_COUNT,
                                                    (h.lastCol - h.firstCol + 1) * PIXEL_COUNT, rowSize,
                                                    firstInv - lastInv + 1));
    pmSubscribeToMemory(ti.taskHandle, di.deviceHandle, si.subtaskId, si.splitInfo, OUTPUT_MEM_INDEX, WRITE_SUBSCRIPTION,
                        pmScatteredSubscriptionInfo(sr * rowSize + sc * PIXEL_COUNT, (ec - sc) * PIXEL_COUNT, rowSize,
                                                    er - sr));
    return pmSuccess;
}

// CPU callback (reference imageFiltering.cpp:379-472): char accumulators in mode 0.
static pmStatus imageFilter_cpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si) {
    const imageFilterTaskConf* c = (const imageFilterTaskConf*)ti.taskConf;
    int sc, ec, sr, er;
    if (!GetSubtaskSubscription(c, si.subtaskId, si.splitInfo, &sc, &ec, &sr, &er)) return pmSuccess;
    const Halo h = halo_of(c, sc, ec, sr, er);
    const pmMemInfo& in = si.memInfo[INPUT_MEM_INDEX];
    const pmMemInfo& out = si.memInfo[OUTPUT_MEM_INDEX];
    const char* inv = (const char*)in.ptr;
    char* o = (char*)out.ptr;
    const size_t inWidth = in.visibilityType == SUBSCRIPTION_NATURAL ? c->imageWidth : (size_t)(h.lastCol - h.firstCol + 1);
    const size_t outWidth = out.visibilityType == SUBSCRIPTION_NATURAL ? c->imageWidth : (size_t)(ec - sc);
    const int totalRows = h.lastRow - h.firstRow + 1;
    const int r = (int)c->filterRadius;
    for (int i = sr; i < er; ++i) {
        for (int j = sc; j < ec; ++j) {
            int acc[3] = {0, 0, 0};   // input channels 0, 1, 2
            char cacc[3] = {0, 0, 0};
            for (int k = i - r; k <= i + r; ++k) {
                for (int l = j - r; l <= j + r; ++l) {
                    const int m = k < 0 ? 0 : (k >= (int)c->imageHeight ? (int)c->imageHeight - 1 : k);
                    const int n = l < 0 ? 0 : (l >= (int)c->imageWidth ? (int)c->imageWidth - 1 : l);
                    const size_t idx = ((size_t)(totalRows - 1 - (m - h.firstRow)) * inWidth + (n - h.firstCol)) * PIXEL_COUNT;
                    const char w = c->filter[k - (i - r)][l - (j - r)];
                    if (c->arithmeticMode == 0) {
                        for (int q = 0; q < 3; ++q) cacc[q] = (char)(cacc[q] + inv[idx + q] * w);
                    } else {
                        for (int q = 0; q < 3; ++q) acc[q] += (int)(unsigned char)inv[idx + q] * (int)w;
                    }
                }
            }
            const size_t off = ((size_t)(i - sr) * outWidth + (j - sc)) * PIXEL_COUNT;
            for (int q = 0; q < 3; ++q) {
                if (c->arithmeticMode == 0) {
                    o[off + q] = cacc[2 - q];
                } else {
                    int v = acc[2 - q] >> c->rightShift;
                    o[off + q] = (char)(unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
                }
            }
        }
    }
    return pmSuccess;
}

// GPU callback: describes the subtask's view to the kernel (replaces
// imageFilter_cudaLaunchFunc, imageFilteringCUDA.cu:295-342).
static pmStatus imageFilter_gpu(pmTaskInfo ti, pmDeviceInfo, pmSubtaskInfo si, void* stream) {
    const imageFilterTaskConf* c = (const imageFilterTaskConf*)ti.taskConf;
    int sc, ec, sr, er;
    if (!GetSubtaskSubscription(c, si.subtaskId, si.splitInfo, &sc, &ec, &sr, &er)) return pmSuccess;
    const Halo h = halo_of(c, sc, ec, sr, er);
    const pmMemInfo& in = si.memInfo[INPUT_MEM_INDEX];
    const pmMemInfo& out = si.memInfo[OUTPUT_MEM_INDEX];
    uh_imagefilter_args a;
    memset(&a, 0, sizeof(a));
    a.in = (const uint8_t*)in.ptr;
    a.in_stride = (in.visibilityType == SUBSCRIPTION_NATURAL ? c->imageWidth : (size_t)(h.lastCol - h.firstCol + 1)) * PIXEL_COUNT;
    a.out = (uint8_t*)out.ptr;
    a.out_stride = (out.visibilityType == SUBSCRIPTION_NATURAL ? c->imageWidth : (size_t)(ec - sc)) * PIXEL_COUNT;
    a.row0 = sr;
    a.col0 = sc;
    a.rows = er - sr;
    a.cols = ec - sc;
    a.view_last_row = h.lastRow;
    a.view_first_col = h.firstCol;
    a.height = (int)c->imageHeight;
    a.width = (int)c->imageWidth;
    a.radius = (int)c->filterRadius;
    a.mode = c->arithmeticMode;
    a.shift = c->rightShift;
    const int d = 2 * a.radius + 1;
    for (int y = 0; y < d; ++y)
        for (int x = 0; x < d; ++x) a.weights[y * d + x] = (int8_t)c->filter[y][x];
    if (uh_imagefilter(&a, stream) != 0) {
        apps::set_error(std::string("uh_imagefilter: ") + uh_last_error());
        return pmGraphicsCardError;
    }
    return pmSuccess;
}

struct Instance {
    imageFilterTaskConf conf;
    std::shared_ptr<apps::SharedMem> inputs;   // the input image, shared with clones
    pmMemHandle in = nullptr, out = nullptr;
    size_t bytes = 0;
    unsigned long subtasks = 0;
};

static apps::CallbackVariants g_cb;
static std::mutex g_mtx;

}  // namespace imageFiltering

using namespace imageFiltering;

extern "C" void* ua_if_create(const ua_if_config* cfg) {
    if (!cfg || cfg->width <= 0 || cfg->height <= 0 || cfg->radius < 1 || cfg->radius > 15 || cfg->tile <= 0 ||
        (cfg->mode != 0 && cfg->mode != 1)) {
        apps::set_error("ua_if_create: invalid configuration");
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> g(g_mtx);
        pmCallbacks cb;
        cb.dataDistribution = imageFilterDataDistribution;
        cb.subtask_cpu = imageFilter_cpu;
        cb.subtask_gpu_custom = imageFilter_gpu;
        if (apps::register_variants("IMAGEFILTER", cb, g_cb) != 0) {
            apps::set_error(std::string("register callbacks: ") + pmGetLastError());
            return nullptr;
        }
    }
    Instance* in = new Instance();
    memset(&in->conf, 0, sizeof(in->conf));
    in->conf.imageWidth = (size_t)cfg->width;
    in->conf.imageHeight = (size_t)cfg->height;
    in->conf.imageBytesPerLine = (size_t)cfg->width * PIXEL_COUNT;
    in->conf.filterRadius = (size_t)cfg->radius;
    in->conf.arithmeticMode = cfg->mode;
    in->conf.rightShift = cfg->shift;
    in->conf.tileDim = cfg->tile;
    const int d = 2 * cfg->radius + 1;
    for (int y = 0; y < d; ++y)
        for (int x = 0; x < d; ++x) in->conf.filter[y][x] = (char)cfg->weights[y * d + x];
    in->bytes = (size_t)cfg->width * cfg->height * PIXEL_COUNT;
    const unsigned long tx = (cfg->width + cfg->tile - 1) / cfg->tile, ty = (cfg->height + cfg->tile - 1) / cfg->tile;
    in->subtasks = tx * ty;
    in->inputs = std::make_shared<apps::SharedMem>();
    if (pmCreateMemory2D(cfg->height, (size_t)cfg->width * PIXEL_COUNT, &in->in) != pmSuccess) in->in = nullptr;
    in->inputs->handles.push_back(in->in);
    if (!in->in || pmCreateMemory2D(cfg->height, (size_t)cfg->width * PIXEL_COUNT, &in->out) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    return in;
}

// A second instance over the same input image with its own output.
extern "C" void* ua_if_clone(void* h) {
    const Instance* src = (const Instance*)h;
    if (!src) return nullptr;
    Instance* in = new Instance();
    in->conf = src->conf;
    in->bytes = src->bytes;
    in->subtasks = src->subtasks;
    in->inputs = src->inputs;
    in->in = src->in;
    if (pmCreateMemory2D(src->conf.imageHeight, src->conf.imageWidth * PIXEL_COUNT, &in->out) != pmSuccess) {
        apps::set_error(std::string("create memory: ") + pmGetLastError());
        delete in;
        return nullptr;
    }
    return in;
}

extern "C" int ua_if_set_input(void* h, const uint8_t* image) {
    Instance* in = (Instance*)h;
    pmRawMemPtr p;
    pmStatus s = pmGetRawMemPtr(in->in, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(p, image, in->bytes);
    return 0;
}

extern "C" int ua_if_run(void* h, int device, double* seconds) {
    Instance* in = (Instance*)h;
    if (device < 0 || device > 2) return -1;
    pmTaskMem mem[MAX_MEM_INDICES];
    mem[INPUT_MEM_INDEX] = pmTaskMem(in->in, READ_ONLY, SUBSCRIPTION_OPTIMAL);
    mem[OUTPUT_MEM_INDEX] = pmTaskMem(in->out, WRITE_ONLY, SUBSCRIPTION_OPTIMAL);
    pmTaskDetails d(&in->conf, sizeof(in->conf), mem, MAX_MEM_INDICES, g_cb.h[device], in->subtasks);
    d.canSplitCpuSubtasks = apps::split_request().cpu;
    d.canSplitGpuSubtasks = apps::split_request().gpu;
    d.multiAssignEnabled = apps::split_request().multi_assign;
    d.policy = SLOW_START;
    return apps::run_task(d, seconds);
}

extern "C" int ua_if_get_output(void* h, uint8_t* out) {
    Instance* in = (Instance*)h;
    pmStatus s = pmFetchMemory(in->out);
    if (s != pmSuccess) return (int)s;
    pmRawMemPtr p;
    s = pmGetRawMemPtr(in->out, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(out, p, in->bytes);
    return 0;
}

extern "C" int ua_if_fetch(void* h, uint64_t* checksum) {
    Instance* in = (Instance*)h;
    return apps::fetch_output(in->out, in->bytes, checksum);
}

extern "C" int ua_if_get_input(void* h, uint8_t* image) {
    Instance* in = (Instance*)h;
    if (!in || !image) return -1;
    pmStatus s = pmFetchMemory(in->in);
    pmRawMemPtr p;
    if (s == pmSuccess) s = pmGetRawMemPtr(in->in, &p);
    if (s != pmSuccess) return (int)s;
    memcpy(image, p, in->bytes);
    return 0;
}

extern "C" int ua_if_get_config(void* h, ua_if_config* cfg) {
    const Instance* in = (const Instance*)h;
    if (!in || !cfg) return -1;
    memset(cfg, 0, sizeof(*cfg));
    cfg->width = (int32_t)in->conf.imageWidth;
    cfg->height = (int32_t)in->conf.imageHeight;
    cfg->radius = (int32_t)in->conf.filterRadius;
    cfg->mode = in->conf.arithmeticMode;
    cfg->shift = in->conf.rightShift;
    cfg->tile = (int32_t)in->conf.tileDim;
    const int d = 2 * cfg->radius + 1;
    for (int y = 0; y < d; ++y)
        for (int x = 0; x < d; ++x) cfg->weights[y * d + x] = (int8_t)in->conf.filter[y][x];
    return 0;
}

extern "C" int ua_if_destroy(void* h) {
    Instance* in = (Instance*)h;
    if (!in) return 0;
    pmReleaseMemory(in->out);
    delete in;    // the input goes with the last holder of `inputs`
    return 0;
}
