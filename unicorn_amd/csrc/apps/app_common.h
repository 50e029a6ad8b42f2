// Shared helpers of the benchmark apps: callback registration per device set (the reference's
// callback nulling, testSuite/common/source/code/src/commonAPI.cpp:85-130), timing and status
// plumbing for the C-ABI drivers.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "pmPublicDefinitions.h"
#include "pmPublicUtilities.h"
#include "pmExtensions.h"
#include "unicorn_apps.h"

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

void set_error(const std::string& s);

}  // namespace apps
