// ORACLE — test infrastructure only.
//
// Driver for the reference's own serial functions, compiled from the reference sources by
// oracle/Makefile into oracle/_ref/ref_runner (this container only; the GPU host never builds
// or needs it). It is used to write small golden vectors and to pin the C restatement in
// oracle.c. It calls, unchanged:
//   imageFiltering::serialImageFilter   testSuite/imageFiltering/source/code/src/imageFiltering.cpp:137-211
//   imageFiltering::imageFilter_cpu     imageFiltering.cpp:379-472 (one call per tile, NATURAL views)
//   pageRank::serialPageRank            testSuite/pageRank/source/code/src/pageRank.cpp:237-276 (2 iterations)
//   matrixMultiply::serialMatrixMultiply testSuite/matrixMultiply/source/code/src/matrixMultiply.cpp:47-55 (int)
//   matrixTranspose::serialMatrixTranspose testSuite/matrixTranspose/source/code/src/matrixTranspose.cpp:187-261
//   matrixTranspose::transposeMatrixBlock  matrixTranspose.cpp:90-110 (square blocks, both forms)
// The pm:: structures passed to imageFilter_cpu are laid out from the reference's own header;
// nothing of the reference runtime is called (those references are discarded by --gc-sections).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pmPublicDefinitions.h"

namespace imageFiltering {
extern size_t gImageWidth, gImageHeight;
void serialImageFilter(void* pImageData, size_t pFilterRadius, char* pSerialOutput, char pFilter[32][32]);
pm::pmStatus imageFilter_cpu(pm::pmTaskInfo pTaskInfo, pm::pmDeviceInfo pDeviceInfo, pm::pmSubtaskInfo pSubtaskInfo);
}
namespace pageRank {
extern unsigned int gTotalWebPages;
extern unsigned int gMaxOutlinksPerWebPage;
extern float* gSerialOutput;
void serialPageRank(unsigned int* pWebDump);
}
namespace matrixTranspose {
void serialMatrixTranspose(bool pInplace, int* pInputMatrix, int* pOutputMatrix, size_t pInputDimRows, size_t pInputDimCols);
void transposeMatrixBlock(int* pBlock, size_t pBlockDim);
void transposeMatrixBlock(int* pInputBlock, int* pOutputBlock, size_t pBlockDim);
}
namespace matrixMultiply {
void serialMatrixMultiply(int* pMatrixA, int* pMatrixB, int* pMatrixC, int pDim1, int pDim2, int pDim3);
}

static std::vector<char> slurp(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<char> v((size_t)n);
    if (n && fread(v.data(), 1, (size_t)n, f) != (size_t)n) exit(2);
    fclose(f);
    return v;
}

static void spit(const char* path, const void* p, size_t n) {
    FILE* f = fopen(path, "wb");
    if (!f || fwrite(p, 1, n, f) != n) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fclose(f);
}

// Layout of the reference's imageFilterTaskConf (imageFiltering.h:96-109, default build).
struct RefImageFilterTaskConf {
    size_t imageWidth, imageHeight, imageOffset, imageBytesPerLine, filterRadius;
    char imagePath[256];
    char filter[32][32];
};

template <typename T>
static T& raw_object(std::vector<unsigned char>& storage) {
    storage.assign(sizeof(T), 0);
    return *reinterpret_cast<T*>(storage.data());
}

int main(int argc, char** argv) {
    if (argc < 2) { fprintf(stderr, "usage: ref_runner if|iftile|pr|mm ...\n"); return 2; }
    const char* mode = argv[1];
    if (!strcmp(mode, "if") && argc == 8) {
        // if W H r in.bin filter.bin out.bin   (filter.bin: 32*32 int8, row-major [32][32])
        size_t W = strtoul(argv[2], 0, 10), H = strtoul(argv[3], 0, 10), r = strtoul(argv[4], 0, 10);
        std::vector<char> in = slurp(argv[5]), filt = slurp(argv[6]);
        imageFiltering::gImageWidth = W;
        imageFiltering::gImageHeight = H;
        std::vector<char> out(W * H * 3);
        char f[32][32];
        memcpy(f, filt.data(), sizeof(f));
        imageFiltering::serialImageFilter(in.data(), r, out.data(), f);
        spit(argv[7], out.data(), out.size());
        return 0;
    }
    if (!strcmp(mode, "iftile") && argc == 9) {
        // iftile W H r tile in.bin filter.bin out.bin: imageFilter_cpu on every tile subtask with
        // NATURAL views of the address spaces (the reference CPU stub's view)
        size_t W = strtoul(argv[2], 0, 10), H = strtoul(argv[3], 0, 10), r = strtoul(argv[4], 0, 10);
        size_t tile = strtoul(argv[5], 0, 10);
        if (tile != 2048) { fprintf(stderr, "the reference's TILE_DIM is 2048\n"); return 2; }
        std::vector<char> in = slurp(argv[6]), filt = slurp(argv[7]);
        std::vector<char> out(W * H * 3, 0);
        std::vector<unsigned char> sconf, sti, sdi, ssi;
        RefImageFilterTaskConf& conf = raw_object<RefImageFilterTaskConf>(sconf);
        conf.imageWidth = W;
        conf.imageHeight = H;
        conf.imageBytesPerLine = W * 3;
        conf.filterRadius = r;
        memcpy(conf.filter, filt.data(), sizeof(conf.filter));
        pm::pmTaskInfo& ti = raw_object<pm::pmTaskInfo>(sti);
        ti.taskConf = &conf;
        ti.taskConfLength = sizeof(conf);
        pm::pmDeviceInfo& di = raw_object<pm::pmDeviceInfo>(sdi);
        const size_t tx = (W + tile - 1) / tile, ty = (H + tile - 1) / tile;
        ti.subtaskCount = tx * ty;
        for (size_t s = 0; s < tx * ty; ++s) {
            pm::pmSubtaskInfo& si = raw_object<pm::pmSubtaskInfo>(ssi);
            si.subtaskId = s;
            const long sc = (long)((s % tx) * tile), sr = (long)((s / tx) * tile);
            const long ec = sc + (long)tile < (long)W ? sc + (long)tile : (long)W;
            const long er = sr + (long)tile < (long)H ? sr + (long)tile : (long)H;
            const long fr = sr - (long)r < 0 ? 0 : sr - (long)r;
            const long lr = er - 1 + (long)r >= (long)H ? (long)H - 1 : er - 1 + (long)r;
            const long fc = sc - (long)r < 0 ? 0 : sc - (long)r;
            (void)fr;
            // NATURAL view pointers: consolidated subscription starts
            si.memInfo[0].ptr = in.data() + ((H - 1 - lr) * W + fc) * 3;
            si.memInfo[0].visibilityType = pm::SUBSCRIPTION_NATURAL;
            si.memInfo[1].ptr = out.data() + (sr * W + sc) * 3;
            si.memInfo[1].visibilityType = pm::SUBSCRIPTION_NATURAL;
            si.memCount = 2;
            (void)ec;
            imageFiltering::imageFilter_cpu(ti, di, si);
        }
        spit(argv[8], out.data(), out.size());
        return 0;
    }
    if (!strcmp(mode, "pr") && argc == 6) {
        // pr N maxOut records.bin out.bin: records of (maxOut+1) uint32 per page, 1-based targets
        unsigned int N = (unsigned int)strtoul(argv[2], 0, 10), M = (unsigned int)strtoul(argv[3], 0, 10);
        std::vector<char> rec = slurp(argv[4]);
        if (rec.size() != (size_t)N * (M + 1) * 4) { fprintf(stderr, "bad record file\n"); return 2; }
        pageRank::gTotalWebPages = N;
        pageRank::gMaxOutlinksPerWebPage = M;
        std::vector<float> out(N);
        pageRank::gSerialOutput = out.data();
        pageRank::serialPageRank((unsigned int*)rec.data());
        spit(argv[5], out.data(), out.size() * 4);
        return 0;
    }
    if (!strcmp(mode, "mm") && argc == 6) {
        int n = atoi(argv[2]);
        std::vector<char> a = slurp(argv[3]), b = slurp(argv[4]);
        std::vector<int> c((size_t)n * n, 0);
        matrixMultiply::serialMatrixMultiply((int*)a.data(), (int*)b.data(), c.data(), n, n, n);
        spit(argv[5], c.data(), c.size() * 4);
        return 0;
    }
    if (!strcmp(mode, "mt") && argc == 7) {
        // mt inplace rows cols in.bin out.bin : serialMatrixTranspose (in place: on a copy of in)
        const bool inplace = atoi(argv[2]) != 0;
        size_t rows = strtoul(argv[3], 0, 10), cols = strtoul(argv[4], 0, 10);
        std::vector<char> a = slurp(argv[5]);
        std::vector<int> out(rows * cols, 0);
        if (inplace) {
            memcpy(out.data(), a.data(), out.size() * 4);
            matrixTranspose::serialMatrixTranspose(true, nullptr, out.data(), rows, cols);
        } else {
            matrixTranspose::serialMatrixTranspose(false, (int*)a.data(), out.data(), rows, cols);
        }
        spit(argv[6], out.data(), out.size() * 4);
        return 0;
    }
    if (!strcmp(mode, "mtblock") && argc == 6) {
        // mtblock inplace bs in.bin out.bin : transposeMatrixBlock of one bs x bs block
        const bool inplace = atoi(argv[2]) != 0;
        size_t bs = strtoul(argv[3], 0, 10);
        std::vector<char> a = slurp(argv[4]);
        std::vector<int> out(bs * bs, 0);
        if (inplace) {
            memcpy(out.data(), a.data(), out.size() * 4);
            matrixTranspose::transposeMatrixBlock(out.data(), bs);
        } else {
            matrixTranspose::transposeMatrixBlock((int*)a.data(), out.data(), bs);
        }
        spit(argv[5], out.data(), out.size() * 4);
        return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 2;
}
