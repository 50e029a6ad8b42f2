// Shared helpers of the benchmark apps: callback registration per device set (the reference's
// callback nulling, testSuite/common/source/code/src/commonAPI.cpp:85-130), timing and status
// plumbing for the C-ABI drivers.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "pmPublicDefinitions.h"
#include "pmPublicUtilities.h"
#include "pmExtensions.h"
#include "unicorn_apps.h"
#include "unicorn_hip.h"

#include <cstdlib>

namespace apps {

using namespace pm;

inline double now_secs() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Callback handles of one app for the three device sets.
struct CallbackVariants {
    pmCallbackHandle h[3] = {nullptr, nullptr, nullptr};   // UA_DEVICE_CPU, _GPU, _BOTH
    bool registered = false;
};

// Registers `cb` under key+"_CPU"/"_GPU"/"_BOTH" with the other device kinds nulled.
inline int register_variants(const char* key, pmCallbacks cb, CallbackVariants& v) {
    if (v.registered) return 0;
    pmCallbacks cpu = cb, gpu = cb, both = cb;
    cpu.subtask_gpu_custom = nullptr;
    cpu.subtask_gpu_hip = nullptr;
    gpu.subtask_cpu = nullptr;
    std::string k = key;
    if (pmRegisterCallbacks((k + "_CPU").c_str(), cpu, &v.h[0]) != pmSuccess) return -1;
    if (pmRegisterCallbacks((k + "_GPU").c_str(), gpu, &v.h[1]) != pmSuccess) return -1;
    if (pmRegisterCallbacks((k + "_BOTH").c_str(), both, &v.h[2]) != pmSuccess) return -1;
    v.registered = true;
    return 0;
}

// Subtask splitting requested by the apps that support it (imageFiltering, matrixMultiply, fft),
// as the reference apps set pmTaskDetails::canSplitCpuSubtasks (imageFiltering.cpp:264,
// matrixMultiplyBlas.cpp:219, fft.cpp:435-438). CPU splitting is on by default, GPU splitting off; multi_assign
// is pmTaskDetails::multiAssignEnabled of every app's tasks (default true, as the reference).
struct SplitRequest {
    bool cpu = true;
    bool gpu = false;
    bool multi_assign = true;
};
SplitRequest& split_request();

// Runs one task to completion; returns its pmStatus (0 = success) and the wall time.
inline int run_task(pmTaskDetails& d, double* secs) {
    pmTaskHandle th = nullptr;
    const double t0 = now_secs();
    pmStatus s = pmSubmitTask(d, &th);
    if (s != pmSuccess) return (int)s;
    s = pmWaitForTaskCompletion(th);
    const double t1 = now_secs();
    pmReleaseTask(th);
    if (secs) *secs = t1 - t0;
    return (int)s;
}

// Input address spaces shared by an app instance and its clones (a clone has its own outputs);
// they are released together with the last holder.
struct SharedMem {
    std::vector<pmMemHandle> handles;
    ~SharedMem() {
        for (pmMemHandle h : handles)
            if (h) pmReleaseMemory(h);
    }
};

// 64-bit FNV-1a over 8-byte words (then the tail bytes): a cheap content hash of task outputs.
inline uint64_t hash_bytes(const void* p, size_t n) {
    const uint8_t* b = (const uint8_t*)p;
    uint64_t h = 0xcbf29ce484222325ull;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
    }
    for (; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

// Brings an output address space to the host (pmFetchMemory) and hashes it when asked.
inline int fetch_output(pmMemHandle m, size_t bytes, uint64_t* checksum) {
    pmStatus s = pmFetchMemory(m);
    if (s != pmSuccess) return (int)s;
    if (checksum) {
        pmRawMemPtr p;
        s = pmGetRawMemPtr(m, &p);
        if (s != pmSuccess) return (int)s;
        *checksum = hash_bytes(p, bytes);
    }
    return 0;
}

// Picks the GEMM tile configuration for a split-BF16 product shape by timing (once per shape
// class); UA_GEMM_AUTOTUNE=0 keeps the size-based default.
inline void autotune_gemm(int M, int N, int K, int precision) {
    if (precision != 0 || !pmHasHipDevice()) return;
    const char* e = getenv("UA_GEMM_AUTOTUNE");
    if (e && atoi(e) == 0) return;
    uh_gemm_autotune(M, N, K, 0, 5, 0, nullptr, nullptr);
}

// Tunes on the shape (M, N, K) and records the winner for the shape class of (m, n, k): for
// products that run batched, where the batch rather than one product fills the device.
inline void autotune_gemm_as(int M, int N, int K, int m, int n, int k, int precision) {
    if (precision != 0 || !pmHasHipDevice()) return;
    const char* e = getenv("UA_GEMM_AUTOTUNE");
    if (e && atoi(e) == 0) return;
    int tile = -1;
    if (uh_gemm_autotune(M, N, K, 0, 5, 0, &tile, nullptr) == 0 && tile >= 0) uh_gemm_set_tile(m, n, k, 0, tile);
}

void set_error(const std::string& s);

}  // namespace apps
