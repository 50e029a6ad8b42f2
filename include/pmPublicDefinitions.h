/*
 * unicorn-mi355x — public API of the partitioned-memory task library (namespace pm).
 *
 * Source-compatible with the reference's application interface
 * (source/code/inc/pmPublicDefinitions.h:45-516): same type names, field order, enum
 * values and function signatures, so the benchmark apps compile against either.
 * Differences, all on the GPU side (the reference is CUDA-only there):
 *   - pmDeviceType gains GPU_HIP (the reference's GPU_CUDA slot, pmPublicDefinitions.h:235-242).
 *   - the kernel-kind GPU callback field is subtask_gpu_hip (reference: subtask_gpu_cuda,
 *     pmPublicDefinitions.h:299,314); subtask_gpu_custom keeps its name and receives a
 *     hipStream_t passed as void*.
 *   - pmGpuLaunchConf / pmSetGpuLaunchConf / pmReserveGpuGlobalMem replace the CUDA-named
 *     pmCudaLaunchConf / pmSetCudaLaunchConf / pmReserveCudaGlobalMem (389-418).
 *   - pmTaskDetails.gpuCacheEnabled replaces cudaCacheEnabled (452).
 * Every API call returns pmStatus and never throws (reference pmPublicDefinitions.cpp:50-66).
 */
#ifndef UNICORN_PM_PUBLIC_DEFINITIONS_H
#define UNICORN_PM_PUBLIC_DEFINITIONS_H

#include <stdlib.h>

#if defined(__HIPCC__)
#define PM_HOST_DEVICE __host__ __device__
#else
#define PM_HOST_DEVICE
#endif

namespace pm
{
    typedef unsigned short int ushort;
    typedef unsigned int uint;
    typedef unsigned long ulong;

    const size_t MAX_NAME_STR_LEN = 256;
    const size_t MAX_DESC_STR_LEN = 1024;
    const size_t MAX_CB_KEY_LEN = 128;
    const size_t MAX_MEM_SECTIONS_PER_TASK = 3;

    /* Status codes; values match the reference enum order (pmPublicDefinitions.h:61-89). */
    typedef enum pmStatus
    {
        pmSuccess = 0,
        pmOk,
        pmStatusUnavailable,
        pmFatalError,
        pmInitializationFailure,
        pmNetworkInitError,
        pmNetworkTerminationError,
        pmInvalidIndex,
        pmInvalidCommand,
        pmThreadingLibraryFailure,
        pmTimerFailure,
        pmMemoryError,
        pmNetworkError,
        pmIgnorableError,
        pmGraphicsCardError,
        pmBeyondComputationalLimits,
        pmUnrecognizedMemory,
        pmInvalidKey,
        pmMaxKeyLengthExceeded,
        pmDataProcessingFailure,
        pmNoCompatibleDevice,
        pmConfFileNotFound,
        pmInvalidOffset,
        pmInvalidCallbacks,
        pmUserError,
        pmMaxStatusValues
    } pmStatus;

    const char* pmGetLibVersion();
    const char* pmGetLastError();
    pmStatus pmInitialize();
    pmStatus pmFinalize();
    uint pmGetHostId();
    uint pmGetHostCount();

    typedef void* pmMemHandle;
    typedef void* pmRawMemPtr;
    typedef void* pmTaskHandle;
    typedef void* pmDeviceHandle;
    typedef void* pmCallbackHandle;
    typedef void* pmClusterHandle;

    typedef enum pmMemType
    {
        READ_ONLY,
        WRITE_ONLY,
        READ_WRITE,
        READ_ONLY_LAZY,
        WRITE_ONLY_LAZY,
        READ_WRITE_LAZY,
        MAX_MEM_TYPE
    } pmMemType;

    typedef enum pmSubscriptionType
    {
        READ_SUBSCRIPTION,
        WRITE_SUBSCRIPTION,
        READ_WRITE_SUBSCRIPTION
    } pmSubscriptionType;

    /* NATURAL: subscribed bytes sit at their address-space distances.
     * COMPACT: disjoint subscribed records are packed back to back in address order.
     * OPTIMAL: resolved per device by the library. */
    typedef enum pmSubscriptionVisibilityType
    {
        SUBSCRIPTION_NATURAL,
        SUBSCRIPTION_COMPACT,
        SUBSCRIPTION_OPTIMAL,
        MAX_SUBSCRIPTION_VISBILITY_TYPE
    } pmSubscriptionVisibilityType;

    typedef struct pmSubscriptionInfo
    {
        size_t offset;
        size_t length;

        pmSubscriptionInfo();
        pmSubscriptionInfo(size_t, size_t);
    } pmSubscriptionInfo;

    typedef struct pmScatteredSubscriptionInfo
    {
        size_t offset;
        size_t size;
        size_t step;
        size_t count;

        pmScatteredSubscriptionInfo();
        pmScatteredSubscriptionInfo(size_t, size_t, size_t, size_t);
    } pmScatteredSubscriptionInfo;

    typedef struct pmGpuContext
    {
        void* scratchBuffer;
        void* reservedGlobalMem;

        pmGpuContext();
    } pmGpuContext;

    typedef struct pmSplitInfo
    {
        uint splitId;
        uint splitCount;

        pmSplitInfo();
        pmSplitInfo(uint, uint);
    } pmSplitInfo;

    typedef struct pmMemInfo
    {
        pmRawMemPtr ptr;
        pmRawMemPtr readPtr;
        pmRawMemPtr writePtr;
        size_t length;
        pmSubscriptionVisibilityType visibilityType;

        pmMemInfo();
        pmMemInfo(pmRawMemPtr, pmRawMemPtr, pmRawMemPtr, size_t);
        pmMemInfo(pmRawMemPtr, pmRawMemPtr, pmRawMemPtr, size_t, pmSubscriptionVisibilityType);
    } pmMemInfo;

    typedef struct pmSubtaskInfo
    {
        ulong subtaskId;
        pmMemInfo memInfo[MAX_MEM_SECTIONS_PER_TASK];
        uint memCount;
        pmGpuContext gpuContext;
        pmSplitInfo splitInfo;

        pmSubtaskInfo();
        pmSubtaskInfo(ulong, pmMemInfo*, uint);
    } pmSubtaskInfo;

    typedef struct pmTaskInfo
    {
        pmTaskHandle taskHandle;
        void* taskConf;
        uint taskConfLength;
        ulong taskId;
        ulong subtaskCount;
        ushort priority;
        uint originatingHost;

        pmTaskInfo();
    } pmTaskInfo;

    typedef struct pmDataTransferInfo
    {
        pmMemHandle memHandle;
        size_t memLength;
        size_t* operatedMemLength;
        pmMemType memType;
        uint srcHost;
        uint destHost;

        pmDataTransferInfo();
    } pmDataTransferInfo;

    typedef enum pmDeviceType
    {
        CPU = 0,
        GPU_HIP,
        MAX_DEVICE_TYPES
    } pmDeviceType;

    typedef struct pmDeviceInfo
    {
        pmDeviceHandle deviceHandle;
        char name[MAX_NAME_STR_LEN];
        char description[MAX_DESC_STR_LEN];
        pmDeviceType deviceType;
        uint host;
        uint deviceIdOnHost;
        uint deviceIdInCluster;

        pmDeviceInfo();
    } pmDeviceInfo;

    typedef enum pmSchedulingPolicy
    {
        SLOW_START,
        RANDOM_STEAL,
        RANDOM_STEAL_WITH_AFFINITY,
        EQUAL_STATIC,
        PROPORTIONAL_STATIC,
        NODE_EQUAL_STATIC
    } pmSchedulingPolicy;

    typedef enum pmAffinityCriterion
    {
        MAXIMIZE_LOCAL_DATA,
        MINIMIZE_REMOTE_SOURCES,
        MINIMIZE_REMOTE_TRANSFER_EVENTS,
        MINIMIZE_REMOTE_TRANSFERS_ESTIMATED_TIME,
        DERIVED_AFFINITY,
        MAX_AFFINITY_CRITERION
    } pmAffinityCriterion;

    typedef enum pmScratchBufferType
    {
        PRE_SUBTASK_TO_SUBTASK,
        SUBTASK_TO_POST_SUBTASK,
        PRE_SUBTASK_TO_POST_SUBTASK,
        REDUCTION_TO_REDUCTION
    } pmScratchBufferType;

    typedef struct pmRedistributionMetadata
    {
        uint order;
        uint count;

        pmRedistributionMetadata();
        pmRedistributionMetadata(uint, uint);
    } pmRedistributionMetadata;

    /* Application callbacks (reference pmPublicDefinitions.h:297-306). */
    typedef pmStatus (*pmDataDistributionCallback)(pmTaskInfo pTaskInfo, pmDeviceInfo pDeviceInfo, pmSubtaskInfo pSubtaskInfo);
    typedef pmStatus (*pmSubtaskCallback_CPU)(pmTaskInfo pTaskInfo, pmDeviceInfo pDeviceInfo, pmSubtaskInfo pSubtaskInfo);
    /* A __global__ kernel; the library launches it with the subtask's pmGpuLaunchConf on the
     * subtask stream. taskInfo.taskConf and pDeviceInfo are device copies. */
    typedef void (*pmSubtaskCallback_GPU_HIP)(pmTaskInfo pTaskInfo, pmDeviceInfo* pDeviceInfo, pmSubtaskInfo pSubtaskInfo, pmStatus* pStatus);
    /* A host function that enqueues work on pStream (a hipStream_t) and must not synchronise. */
    typedef pmStatus (*pmSubtaskCallback_GPU_Custom)(pmTaskInfo pTaskInfo, pmDeviceInfo pDeviceInfo, pmSubtaskInfo pSubtaskInfo, void* pStream);
    typedef pmStatus (*pmDataReductionCallback)(pmTaskInfo pTaskInfo, pmDeviceInfo pDevice1Info, pmSubtaskInfo pSubtask1Info, pmDeviceInfo pDevice2Info, pmSubtaskInfo pSubtask2Info);
    typedef pmStatus (*pmDataRedistributionCallback)(pmTaskInfo pTaskInfo, pmDeviceInfo pDeviceInfo, pmSubtaskInfo pSubtaskInfo);
    typedef bool     (*pmDeviceSelectionCallback)(pmTaskInfo pTaskInfo, pmDeviceInfo pDeviceInfo);
    typedef pmStatus (*pmPreDataTransferCallback)(pmTaskInfo pTaskInfo, pmDataTransferInfo pDataTransferInfo);
    typedef pmStatus (*pmPostDataTransferCallback)(pmTaskInfo pTaskInfo, pmDataTransferInfo pDataTransferInfo);
    typedef pmStatus (*pmTaskCompletionCallback)(pmTaskInfo pTaskInfo);

    typedef struct pmCallbacks
    {
        public:
            pmDataDistributionCallback dataDistribution;
            pmSubtaskCallback_CPU subtask_cpu;
            pmSubtaskCallback_GPU_HIP subtask_gpu_hip;
            pmSubtaskCallback_GPU_Custom subtask_gpu_custom;    /* at most one of the two GPU kinds may be set */
            pmDataReductionCallback dataReduction;
            pmDataRedistributionCallback dataRedistribution;
            pmDeviceSelectionCallback deviceSelection;
            pmPreDataTransferCallback preDataTransfer;
            pmPostDataTransferCallback postDataTransfer;
            pmTaskCompletionCallback taskCompletionCallback;
            const char* subtask_opencl;

            pmCallbacks();
            pmCallbacks(pmDataDistributionCallback, const char* subtask_opencl);
            pmCallbacks(pmDataDistributionCallback, const char* subtask_opencl, pmDataReductionCallback);
            pmCallbacks(pmDataDistributionCallback, const char* subtask_opencl, pmDataRedistributionCallback);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_HIP);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_Custom);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_HIP, pmDataReductionCallback);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_Custom, pmDataReductionCallback);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_HIP, pmDataRedistributionCallback);
            pmCallbacks(pmDataDistributionCallback, pmSubtaskCallback_CPU, pmSubtaskCallback_GPU_Custom, pmDataRedistributionCallback);
    } pmCallbacks;

    pmStatus pmRegisterCallbacks(const char* pKey, pmCallbacks pCallbacks, pmCallbackHandle* pCallbackHandle);
    pmStatus pmReleaseCallbacks(pmCallbackHandle pCallbackHandle);

    pmStatus pmCreateMemory(size_t pLength, pmMemHandle* pMemHandle);
    pmStatus pmCreateMemory2D(size_t pRows, size_t pCols, pmMemHandle* pMemHandle);
    pmStatus pmReleaseMemory(pmMemHandle pMemHandle);
    pmStatus pmFetchMemory(pmMemHandle pMemHandle);
    pmStatus pmFetchMemoryRange(pmMemHandle pMemHandle, size_t pOffset, size_t pLength);
    pmStatus pmGetRawMemPtr(pmMemHandle pMemHandle, pmRawMemPtr* pPtr);

    /* Only callable from a DataDistribution callback (reference pmPublicDefinitions.h:370-377). */
    pmStatus pmSubscribeToMemory(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, uint pMemIndex, pmSubscriptionType pSubscriptionType, const pmSubscriptionInfo& pSubscriptionInfo);
    pmStatus pmSubscribeToMemory(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, uint pMemIndex, pmSubscriptionType pSubscriptionType, const pmScatteredSubscriptionInfo& pScatteredSubscriptionInfo);

    pmStatus pmRedistributeData(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, uint pMemIndex, size_t pOffset, size_t pLength, uint pOrder);

    /* Launch geometry for subtask_gpu_hip (reference pmCudaLaunchConf, pmPublicDefinitions.h:389-402). */
    typedef struct pmGpuLaunchConf
    {
        int blocksX;
        int blocksY;
        int blocksZ;
        int threadsX;
        int threadsY;
        int threadsZ;
        int sharedMem;

        pmGpuLaunchConf();
        pmGpuLaunchConf(int, int, int, int, int, int);
        pmGpuLaunchConf(int, int, int, int, int, int, int);
    } pmGpuLaunchConf;

    pmStatus pmSetGpuLaunchConf(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, pmGpuLaunchConf& pGpuLaunchConf);

    /* Ask the library for a device buffer, visible as pSubtaskInfo.gpuContext.reservedGlobalMem. */
    pmStatus pmReserveGpuGlobalMem(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, size_t pSize);

    typedef struct pmTaskMem
    {
        pmMemHandle memHandle;
        pmMemType memType;
        pmSubscriptionVisibilityType subscriptionVisibilityType;
        bool disjointReadWritesAcrossSubtasks;

        pmTaskMem();
        pmTaskMem(pmMemHandle, pmMemType);
        pmTaskMem(pmMemHandle, pmMemType, pmSubscriptionVisibilityType);
        pmTaskMem(pmMemHandle, pmMemType, pmSubscriptionVisibilityType, bool);
    } pmTaskMem;

    typedef struct pmTaskDetails
    {
        void* taskConf;
        uint taskConfLength;
        pmTaskMem* taskMem;
        uint taskMemCount;
        pmCallbackHandle callbackHandle;
        ulong subtaskCount;
        ulong taskId;
        ushort priority;                        /* default 0 = highest */
        pmSchedulingPolicy policy;              /* default RANDOM_STEAL */
        int timeOutInSecs;                      /* default INT_MAX; negative = none */
        bool multiAssignEnabled;
        bool overlapComputeCommunication;
        bool canSplitCpuSubtasks;
        bool canSplitGpuSubtasks;
        bool gpuCacheEnabled;                   /* default true */
        bool suppressTaskLogs;
        pmAffinityCriterion affinityCriterion;
        pmClusterHandle cluster;

        pmTaskDetails();
        pmTaskDetails(void* taskConf, uint taskConfLength, pmTaskMem*, uint taskMemCount, pmCallbackHandle callbackHandle, ulong subtaskCount);
    } pmTaskDetails;

    pmStatus pmSubmitTask(pmTaskDetails pTaskDetails, pmTaskHandle* pTaskHandle);
    pmStatus pmReleaseTask(pmTaskHandle pTaskHandle);
    pmStatus pmWaitForTaskCompletion(pmTaskHandle pTaskHandle);
    pmStatus pmGetTaskExecutionTimeInSecs(pmTaskHandle pTaskHandle, double* pTime);
    pmStatus pmReleaseTaskAndResources(pmTaskDetails pTaskDetails, pmTaskHandle pTaskHandle);

    /* Scratch buffer shared by the callbacks of one subtask. On the device it is the
     * gpuContext.scratchBuffer of the subtask (reference pmPublicDefinitions.h:493-504). */
    PM_HOST_DEVICE
    inline void* pmGetScratchBuffer(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, pmScratchBufferType pScratchBufferType, size_t pBufferSize, pmGpuContext* pGpuContext)
    {
    #if defined(__HIP_DEVICE_COMPILE__)
        return (pGpuContext ? pGpuContext->scratchBuffer : NULL);
    #else
        void* pmGetScratchBufferHostFunc(pmTaskHandle, pmDeviceHandle, ulong, pmSplitInfo&, pmScratchBufferType, size_t);
        return pmGetScratchBufferHostFunc(pTaskHandle, pDeviceHandle, pSubtaskId, pSplitInfo, pScratchBufferType, pBufferSize);
    #endif
    }

    pmStatus pmReleaseScratchBuffer(pmTaskHandle pTaskHandle, pmDeviceHandle pDeviceHandle, ulong pSubtaskId, pmSplitInfo& pSplitInfo, pmScratchBufferType pScratchBufferType);
    void* pmGetLastReductionScratchBuffer(pmTaskHandle pTaskHandle);
    pmRedistributionMetadata* pmGetRedistributionMetadata(pmTaskHandle pTaskHandle, uint pMemIndex, ulong* pCount);

} // namespace pm

#endif
