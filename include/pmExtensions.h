/*
 * unicorn-mi355x — runtime controls with no counterpart in the reference API.
 *
 * The reference fixes its GPU stub behaviour at compile time (pmInternalDefinitions.h:117-224)
 * and through environment variables read by its test harness (commonAPI.cpp:580-625). This
 * build exposes the HIP stub's execution modes, its HBM budget and its counters here so that
 * drivers and tests can compare them on the same tasks.
 *
 * HIP stub modes:
 *   PM_HIP_MODE_PLAIN     per-subtask stream, launch and staging, as the reference stub:
 *                         subscribed records are gathered on the host into pinned memory,
 *                         copied H2D into a COMPACT device buffer, the GPU callback runs, and
 *                         written records go D2H and are scattered back (pmExecutionStub.cpp
 *                         :3216-3988). Read-only buffers are cached by exact subscription.
 *   PM_HIP_MODE_RESIDENT  address spaces stay mirrored in HBM across tasks; SUBSCRIPTION_OPTIMAL
 *                         resolves to a zero-copy NATURAL view of the mirror (or a device-side
 *                         gather when COMPACT is forced); one launch per subtask.
 *   PM_HIP_MODE_BATCHED   RESIDENT, and the launches of many subtasks (of one or several tasks)
 *                         are recorded into a device work queue and executed by one persistent
 *                         kernel per kernel family, optionally replayed from a hipGraph.
 */
#ifndef UNICORN_PM_EXTENSIONS_H
#define UNICORN_PM_EXTENSIONS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include "pmPublicDefinitions.h"
#endif

#define PM_HIP_MODE_PLAIN 0
#define PM_HIP_MODE_RESIDENT 1
#define PM_HIP_MODE_BATCHED 2

#ifdef __cplusplus
namespace pm
{
#endif

typedef struct pmHipStubStats
{
    uint64_t subtasks;          /* GPU subtasks executed */
    uint64_t batches;           /* work-queue flushes (BATCHED) */
    uint64_t h2d_bytes;         /* host-to-device staging bytes */
    uint64_t d2h_bytes;         /* device-to-host staging bytes */
    uint64_t d2d_bytes;         /* device gather/scatter bytes (COMPACT views of mirrors) */
    uint64_t cache_hits;        /* PLAIN: read-only device buffers reused */
    uint64_t evictions;         /* HBM mirrors evicted to honour the budget */
    /* host-side phase times in nanoseconds, summed */
    uint64_t ns_queue;          /* submit -> first subtask picked by a stub, per task */
    uint64_t ns_stage;          /* BATCHED: distribution + GPU callbacks of a batch's subtasks */
    uint64_t ns_flush;          /* BATCHED: flush host time (queue build, upload, launch) */
    uint64_t ns_wait;           /* BATCHED: stub thread blocked on the oldest batch */
    uint64_t ns_retire;         /* BATCHED: retiring completed batches (scatter, completion) */
    uint64_t ns_wake;           /* task completion -> blocked waiter resumed, per wait */
    uint64_t ma_copies;         /* multi-assigned copies of CPU subtasks run by the HIP stub */
    uint64_t ma_wins;           /* of those, copies that finished first and committed */
    uint64_t oom_waits;         /* device allocations retried after retiring the oldest in-flight subtask */
    uint64_t sentinel_compressed; /* PLAIN: reduction outputs sent back in sentinel-compressed form */
} pmHipStubStats;

#ifdef __cplusplus
/* Options of one submission (pmSubmitTaskWithOptions).
 *   hipMode    PM_HIP_MODE_* for this task only; -1 = the mode set with pmSetHipStubMode.
 *   unordered  start without waiting for earlier tasks on the same address spaces. By default a
 *              task waits for every earlier, unfinished task that it conflicts with on some
 *              address space (one of the two writes it), as the reference's address-space locks
 *              order tasks; an unordered task is for callers that know the two touch disjoint
 *              bytes (e.g. blocks of one matrix). Later tasks still wait for it. */
typedef struct pmTaskSubmitOptions
{
    int hipMode;
    bool unordered;
} pmTaskSubmitOptions;
pmStatus pmSubmitTaskWithOptions(pmTaskDetails pTaskDetails, const pmTaskSubmitOptions* pOptions,
                                 pmTaskHandle* pTaskHandle);
/* Pieces per split subtask for tasks submitted afterwards with canSplitCpuSubtasks or
 * canSplitGpuSubtasks; 0 (the default) takes one piece per stub of the splitting device type, as
 * the reference's split groups do (pmSubtaskSplitter.cpp:49-140), so that with one GPU, GPU
 * splitting is off unless a count is set here. A split subtask's data distribution and subtask
 * callbacks run once per piece with pSubtaskInfo.splitInfo = {splitId, splitCount}; the subtask
 * completes with its last piece. */
pmStatus pmSetSplitFactors(int pCpuPieces, int pGpuPieces);
/* Select the HIP stub mode for tasks submitted afterwards. */
pmStatus pmSetHipStubMode(int pMode);
int pmGetHipStubMode();
/* Force the view used for SUBSCRIPTION_OPTIMAL on the HIP stub in RESIDENT/BATCHED modes:
 * SUBSCRIPTION_NATURAL (default, zero-copy) or SUBSCRIPTION_COMPACT (device gather/scatter). */
pmStatus pmSetHipOptimalView(pmSubscriptionVisibilityType pType);
/* Use hipGraph replay for batch flushes that launch more than one kernel (default on); a
 * one-kernel flush launches directly either way. */
pmStatus pmSetHipGraphs(bool pEnable);
/* Maximum subtasks in flight on the HIP stub (PLAIN/RESIDENT) and per batch (BATCHED). */
pmStatus pmSetHipPipelineDepth(int pDepth);
pmStatus pmSetHipBatchLimit(int pSubtasks);
/* Bytes of HBM the runtime may hold for mirrors and subtask buffers (0 = the default chosen at
 * initialisation: PMLIB_HBM_BUDGET_MB, else 85 % of the free HBM). GPU-resident tasks are admitted
 * in order while their address spaces fit; mirrors no admitted task uses are evicted LRU. */
pmStatus pmSetHbmBudget(size_t pBytes);
size_t pmGetHbmBudget();
size_t pmGetHbmResident();
/* Number of CPU stub threads (0 = one per host core). Takes effect at pmInitialize. */
pmStatus pmSetCpuStubCount(int pCount);
pmStatus pmGetHipStubStats(pmHipStubStats* pStats);
/* Size above which a PLAIN-mode reduction subtask that writes its task's only address space
 * whole sends the output back sentinel-compressed (its non-zero elements and their indices) when
 * fewer than 40 % of the elements are non-zero. The default is the reference's 1 GiB
 * (CUDA_SENTINEL_COMPRESSION_THRESHOLD, pmInternalDefinitions.h:223); a smaller value lets tests
 * and smaller outputs use the path. pPrevious (optional) receives the value replaced. */
pmStatus pmSetSentinelCompressionThreshold(size_t pBytes, size_t* pPrevious);
pmStatus pmResetHipStubStats();
/* Make the HBM mirror of an address space current and keep it (prefetch). */
pmStatus pmPrefetchToDevice(pmMemHandle pMemHandle);
/* Device pointer of an address space's HBM mirror (prefetching it if needed). The caller may
 * read it on the GPU; writes through it must be followed by pmMarkDeviceWritten. */
pmStatus pmGetDeviceMemPtr(pmMemHandle pMemHandle, void** pDevicePtr);
pmStatus pmMarkDeviceWritten(pmMemHandle pMemHandle);
/* Copies bytes [pOffset, pOffset + pLength) of an address space's current contents (from the HBM
 * mirror when that is the newer copy) to pDst without changing where the space is valid: unlike
 * pmFetchMemory + pmGetRawMemPtr, the HBM mirror stays current for the next task. */
pmStatus pmReadMemoryRange(pmMemHandle pMemHandle, size_t pOffset, size_t pLength, void* pDst);
/* Device time of the BATCHED mode's flushes (their persistent kernels), measured with events on
 * the stream the persistent kernels run on: around every flush, or with pmSetHipKernelTimingEvery
 * around every pEvery-th one (0 turns timing off). Each timed flush costs a few microseconds of
 * device time for the two event records, so a sample keeps the measurement from slowing the
 * tasks it measures. pmGetHipKernelTime returns the time and the number of flushes timed. */
pmStatus pmSetHipKernelTiming(bool pEnable);
pmStatus pmSetHipKernelTimingEvery(unsigned int pEvery);
pmStatus pmGetHipKernelTime(double* pMilliseconds, uint64_t* pFlushes, bool pReset);
/* The timed flushes that launched one kernel family only (family numbers as UH_FAMILY_* of
 * unicorn_hip.h): milliseconds, count, and the family's algorithmic work over those flushes
 * (GEMM: flop issued; filter: bytes moved; others: 0). */
pmStatus pmGetHipKernelTimeByFamily(int pFamily, double* pMilliseconds, uint64_t* pFlushes, double* pWork,
                                    bool pReset);
/* Whether a HIP device is available to the runtime. */
bool pmHasHipDevice();

} /* namespace pm */
#endif

#endif
