// Built-in subtask reductions (reference pmPublicUtilities.h:45-75, pmReducer.cpp:593-710,
// pmPublicDefinitions.cpp:750-846). The callback returned by pmGetSubtaskReductionCallbackImpl
// is a marker: the runtime recognises it and folds the subtasks' private outputs itself, on the
// host, or in HBM (uh_reduce) when the HIP stub keeps the shadows there. Both folds take their
// element operations from reduce_ops.h.
#include "pm_internal.h"
#include "../hip/reduce_ops.h"

#include <cstring>

namespace pm {

namespace {

template <pmReductionOpType OP, pmReductionDataType DT>
pmStatus builtin_reduction(pmTaskInfo, pmDeviceInfo, pmSubtaskInfo, pmDeviceInfo, pmSubtaskInfo) {
    // Never called: the runtime performs built-in reductions directly.
    return pmSuccess;
}

template <pmReductionOpType OP>
pmDataReductionCallback pick_type(pmReductionDataType dt) {
    switch (dt) {
        case REDUCE_INTS: return builtin_reduction<OP, REDUCE_INTS>;
        case REDUCE_UNSIGNED_INTS: return builtin_reduction<OP, REDUCE_UNSIGNED_INTS>;
        case REDUCE_LONGS: return builtin_reduction<OP, REDUCE_LONGS>;
        case REDUCE_UNSIGNED_LONGS: return builtin_reduction<OP, REDUCE_UNSIGNED_LONGS>;
        case REDUCE_FLOATS: return builtin_reduction<OP, REDUCE_FLOATS>;
        case REDUCE_DOUBLES: return builtin_reduction<OP, REDUCE_DOUBLES>;
        default: return nullptr;
    }
}

template <int OP, typename T>
void fold(void* dst, const void* src, size_t bytes) {
    T* d = (T*)dst;
    const T* s = (const T*)src;
    const size_t n = bytes / sizeof(T);
    for (size_t i = 0; i < n; ++i) d[i] = redop::apply<OP, T>(d[i], s[i]);
}

template <typename T>
void reduce_typed(void* dst, const void* src, size_t bytes, pmReductionOpType op) {
    switch ((int)op) {
        case redop::ADD: fold<redop::ADD, T>(dst, src, bytes); break;
        case redop::MIN: fold<redop::MIN, T>(dst, src, bytes); break;
        case redop::MAX: fold<redop::MAX, T>(dst, src, bytes); break;
        case redop::PRODUCT: fold<redop::PRODUCT, T>(dst, src, bytes); break;
        case redop::LOGICAL_AND: fold<redop::LOGICAL_AND, T>(dst, src, bytes); break;
        case redop::BITWISE_AND: fold<redop::BITWISE_AND, T>(dst, src, bytes); break;
        case redop::LOGICAL_OR: fold<redop::LOGICAL_OR, T>(dst, src, bytes); break;
        case redop::BITWISE_OR: fold<redop::BITWISE_OR, T>(dst, src, bytes); break;
        case redop::LOGICAL_XOR: fold<redop::LOGICAL_XOR, T>(dst, src, bytes); break;
        case redop::BITWISE_XOR: fold<redop::BITWISE_XOR, T>(dst, src, bytes); break;
        default: rt::raise(pmInvalidCommand, "unknown reduction");
    }
}

}  // namespace

pmDataReductionCallback pmGetSubtaskReductionCallbackImpl(pmReductionOpType op, pmReductionDataType dt) {
    switch (op) {
        case REDUCE_ADD: return pick_type<REDUCE_ADD>(dt);
        case REDUCE_MIN: return pick_type<REDUCE_MIN>(dt);
        case REDUCE_MAX: return pick_type<REDUCE_MAX>(dt);
        case REDUCE_PRODUCT: return pick_type<REDUCE_PRODUCT>(dt);
        case REDUCE_LOGICAL_AND: return pick_type<REDUCE_LOGICAL_AND>(dt);
        case REDUCE_BITWISE_AND: return pick_type<REDUCE_BITWISE_AND>(dt);
        case REDUCE_LOGICAL_OR: return pick_type<REDUCE_LOGICAL_OR>(dt);
        case REDUCE_BITWISE_OR: return pick_type<REDUCE_BITWISE_OR>(dt);
        case REDUCE_LOGICAL_XOR: return pick_type<REDUCE_LOGICAL_XOR>(dt);
        case REDUCE_BITWISE_XOR: return pick_type<REDUCE_BITWISE_XOR>(dt);
        default: return nullptr;
    }
}

pmStatus pmReduceSubtasks(pmTaskHandle, pmDeviceHandle, unsigned long, pmSplitInfo&, pmDeviceHandle,
                          unsigned long, pmSplitInfo&, pmReductionOpType, pmReductionDataType) {
    // Custom reduction callbacks receive NATURAL views of both subtasks' shadows and reduce
    // them directly; the pairwise helper of the reference is not needed on one node.
    rt::set_last_error("pmReduceSubtasks: call reduce on the pmSubtaskInfo views instead");
    return pmFatalError;
}

namespace rt {

bool is_builtin_reduction(pmDataReductionCallback cb, pmReductionOpType* op, pmReductionDataType* dt) {
    for (int o = 0; o < MAX_REDUCTION_OP_TYPES; ++o)
        for (int d = 0; d < MAX_REDUCTION_DATA_TYPES; ++d)
            if (cb && cb == pmGetSubtaskReductionCallbackImpl((pmReductionOpType)o, (pmReductionDataType)d)) {
                *op = (pmReductionOpType)o;
                *dt = (pmReductionDataType)d;
                return true;
            }
    return false;
}

void reduce_host(void* dst, const void* src, size_t bytes, pmReductionOpType op, pmReductionDataType dt) {
    switch (dt) {
        case REDUCE_INTS: reduce_typed<int>(dst, src, bytes, op); break;
        case REDUCE_UNSIGNED_INTS: reduce_typed<unsigned int>(dst, src, bytes, op); break;
        case REDUCE_LONGS: reduce_typed<long>(dst, src, bytes, op); break;
        case REDUCE_UNSIGNED_LONGS: reduce_typed<unsigned long>(dst, src, bytes, op); break;
        case REDUCE_FLOATS: reduce_typed<float>(dst, src, bytes, op); break;
        case REDUCE_DOUBLES: reduce_typed<double>(dst, src, bytes, op); break;
        default: raise(pmInvalidCommand, "unknown reduction data type");
    }
}

}  // namespace rt
}  // namespace pm
