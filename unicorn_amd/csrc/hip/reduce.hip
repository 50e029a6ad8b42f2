// Built-in subtask reductions folded in HBM: dst[i] = dst[i] (op) src[i] for the reference's ten
// operations over its six element types (pmReducer.cpp:593-703; element operations in
// reduce_ops.h, shared with the host fold). A pure streaming kernel: 3 x the bytes of one operand
// move through HBM (two reads and one write), so it is bound by HBM bandwidth; each thread handles
// 16-byte vectors, and the grid is capped at 8 workgroups per CU with a grid-stride loop.
#include "uh_internal.h"
#include "reduce_ops.h"

namespace uh {

template <int OP, typename T>
__global__ void __launch_bounds__(256) k_reduce(T* dst, const T* src, size_t n) {
    constexpr int V = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(V)));
    const size_t nv = n / V;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    vec_t* dv = (vec_t*)dst;
    const vec_t* sv = (const vec_t*)src;
    for (size_t k = i; k < nv; k += stride) {
        vec_t a = dv[k];
        const vec_t b = sv[k];
#pragma unroll
        for (int e = 0; e < V; ++e) a[e] = redop::apply<OP, T>(a[e], b[e]);
        dv[k] = a;
    }
    for (size_t k = nv * V + i; k < n; k += stride) dst[k] = redop::apply<OP, T>(dst[k], src[k]);
}

template <typename T>
static int launch_reduce(int op, void* dst, const void* src, size_t n, hipStream_t s) {
    size_t blocks = (n * sizeof(T) / 16 + 255) / 256;
    const size_t cap = (size_t)cu_count() * 8;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    T* d = (T*)dst;
    const T* a = (const T*)src;
    const dim3 g((unsigned)blocks), b(256);
    switch (op) {
        case redop::ADD: UH_LAUNCH((k_reduce<redop::ADD, T>), g, b, 0, s, d, a, n); break;
        case redop::MIN: UH_LAUNCH((k_reduce<redop::MIN, T>), g, b, 0, s, d, a, n); break;
        case redop::MAX: UH_LAUNCH((k_reduce<redop::MAX, T>), g, b, 0, s, d, a, n); break;
        case redop::PRODUCT: UH_LAUNCH((k_reduce<redop::PRODUCT, T>), g, b, 0, s, d, a, n); break;
        case redop::LOGICAL_AND: UH_LAUNCH((k_reduce<redop::LOGICAL_AND, T>), g, b, 0, s, d, a, n); break;
        case redop::BITWISE_AND: UH_LAUNCH((k_reduce<redop::BITWISE_AND, T>), g, b, 0, s, d, a, n); break;
        case redop::LOGICAL_OR: UH_LAUNCH((k_reduce<redop::LOGICAL_OR, T>), g, b, 0, s, d, a, n); break;
        case redop::BITWISE_OR: UH_LAUNCH((k_reduce<redop::BITWISE_OR, T>), g, b, 0, s, d, a, n); break;
        case redop::LOGICAL_XOR: UH_LAUNCH((k_reduce<redop::LOGICAL_XOR, T>), g, b, 0, s, d, a, n); break;
        case redop::BITWISE_XOR: UH_LAUNCH((k_reduce<redop::BITWISE_XOR, T>), g, b, 0, s, d, a, n); break;
        default: return fail(UH_ERR_ARG, "uh_reduce: unknown operation %d", op);
    }
    UH_LAUNCH_CHECK("k_reduce");
    g_plain_launches++;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_reduce(void* dst, const void* src, size_t count, int op, int type, void* stream) {
    if (!count) return UH_OK;
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    if (!dst || !src) return fail(UH_ERR_ARG, "uh_reduce: null buffer");
    if (((uintptr_t)dst | (uintptr_t)src) & 15u) return fail(UH_ERR_ARG, "uh_reduce: buffers must be 16-byte aligned");
    if (op < 0 || op >= redop::OP_COUNT) return fail(UH_ERR_ARG, "uh_reduce: unknown operation %d", op);
    switch (type) {
        case UH_REDUCE_INT32: return launch_reduce<int32_t>(op, dst, src, count, s);
        case UH_REDUCE_UINT32: return launch_reduce<uint32_t>(op, dst, src, count, s);
        case UH_REDUCE_INT64: return launch_reduce<int64_t>(op, dst, src, count, s);
        case UH_REDUCE_UINT64: return launch_reduce<uint64_t>(op, dst, src, count, s);
        case UH_REDUCE_FLOAT32: return launch_reduce<float>(op, dst, src, count, s);
        case UH_REDUCE_FLOAT64: return launch_reduce<double>(op, dst, src, count, s);
        default: return fail(UH_ERR_ARG, "uh_reduce: unknown element type %d", type);
    }
}

extern "C" int uh_reduce_add_f32(float* dst, const float* src, size_t n, void* stream) {
    return uh_reduce(dst, src, n, redop::ADD, UH_REDUCE_FLOAT32, stream);
}
