/*
 * unicorn-mi355x — public utilities (namespace pm).
 * Same interface as the reference's source/code/inc/pmPublicUtilities.h:43-99:
 * built-in subtask reductions and memory-mapped input files.
 *
 * The built-in reduction returned by pmGetSubtaskReductionCallbackImpl is recognised by the
 * HIP stub, which then reduces the subtasks' private outputs in HBM instead of on the host.
 */
#ifndef UNICORN_PM_PUBLIC_UTILITIES_H
#define UNICORN_PM_PUBLIC_UTILITIES_H

#include "pmPublicDefinitions.h"

namespace pm
{
    typedef enum pmReductionOpType
    {
        REDUCE_ADD,
        REDUCE_MIN,
        REDUCE_MAX,
        REDUCE_PRODUCT,
        REDUCE_LOGICAL_AND,
        REDUCE_BITWISE_AND,
        REDUCE_LOGICAL_OR,
        REDUCE_BITWISE_OR,
        REDUCE_LOGICAL_XOR,
        REDUCE_BITWISE_XOR,
        MAX_REDUCTION_OP_TYPES
    } pmReductionOpType;

    typedef enum pmReductionDataType
    {
        REDUCE_INTS,
        REDUCE_UNSIGNED_INTS,
        REDUCE_LONGS,
        REDUCE_UNSIGNED_LONGS,
        REDUCE_FLOATS,
        REDUCE_DOUBLES,
        MAX_REDUCTION_DATA_TYPES
    } pmReductionDataType;

    pmDataReductionCallback pmGetSubtaskReductionCallbackImpl(pmReductionOpType pOperation, pmReductionDataType pDataType);

    pmStatus pmReduceSubtasks(pmTaskHandle pTaskHandle, pmDeviceHandle pDevice1Handle, unsigned long pSubtask1Id, pmSplitInfo& pSplitInfo1, pmDeviceHandle pDevice2Handle, unsigned long pSubtask2Id, pmSplitInfo& pSplitInfo2, pmReductionOpType pOperation, pmReductionDataType pDataType);

    const size_t MAX_FILE_SIZE_LEN = 2048;

    void* pmGetMappedFile(const char* pPath);
    pmStatus pmMapFile(const char* pPath);
    pmStatus pmUnmapFile(const char* pPath);
    pmStatus pmMapFiles(const char* const* pPaths, uint pFileCount);
    pmStatus pmUnmapFiles(const char* const* pPaths, uint pFileCount);

}   // namespace pm

#endif
