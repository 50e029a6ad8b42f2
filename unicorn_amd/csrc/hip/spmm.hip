// sparseSolver kernel (replaces the cusparseXcoo2csr + cusparseScsrmm2 pair of
// sparseMatrixMultiply_cudaLaunchFunc, reference sparseSolverCUDA.cu:206-290): one block row of
// C = A B with A sparse in COO order (sorted by row, then column) and B, C dense. A's entries are
// 12-byte records {int32 row, int32 col, float value}.
//   C[r][j] = sum over the entries k of row r0 + r, in COO order, of val[k] * B[col[k]][j]
// for r < rows, j < cols; rows without entries become 0.
//
// gfx950 design. One workgroup per output row, 256 threads over the row's columns (up to 8 per
// thread for the app's 2048-wide blocks). The row's entry range comes from two binary searches
// over the sorted row indices (done by one lane, shared through LDS) instead of a CSR conversion
// pass; each entry's B row is read once, coalesced, and accumulated in registers in COO order
// with separately rounded products and sums (no fused multiply-add), so the result equals the
// CPU callback's left-to-right float loop bit for bit.
#include "uh_internal.h"

namespace uh {

struct CooEntry {
    int32_t row, col;
    float val;
};

__device__ __forceinline__ uint32_t lower_row(const CooEntry* e, uint32_t n, int32_t r) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (e[m].row < r) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_spmm_coo(const CooEntry* ent, uint32_t nnz, int32_t row0, const float* B,
                                                   uint64_t ldb, float* C, uint64_t ldc, uint32_t cols) {
    // HIP contracts a * b + c into one fused multiply-add by default (also through __fmul_rn /
    // __fadd_rn, whose header bodies allow it); the CPU callback rounds the product and the sum
    // separately, so contraction is off here and the operators are written out
#pragma clang fp contract(off)
    __shared__ uint32_t range[2];
    const int32_t r = row0 + (int32_t)blockIdx.x;
    if (threadIdx.x == 0) {
        range[0] = lower_row(ent, nnz, r);
        range[1] = lower_row(ent, nnz, r + 1);
    }
    __syncthreads();
    const uint32_t lo = range[0], hi = range[1];
    float* crow = C + (uint64_t)blockIdx.x * ldc;
    for (uint32_t j0 = 0; j0 < cols; j0 += 8 * 256) {
        float acc[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] = 0.f;
        for (uint32_t k = lo; k < hi; ++k) {
            const float a = ent[k].val;
            const float* brow = B + (uint64_t)ent[k].col * ldb;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t j = j0 + threadIdx.x + 256u * e;
                if (j < cols) acc[e] = acc[e] + a * brow[j];
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t j = j0 + threadIdx.x + 256u * e;
            if (j < cols) crow[j] = acc[e];
        }
    }
}

}  // namespace uh

using namespace uh;

extern "C" int uh_spmm_coo_f32(const void* entries, uint32_t nnz, int32_t row0, uint32_t rows, const float* B,
                               uint64_t ldb, float* C, uint64_t ldc, uint32_t cols, void* stream) {
    if (!rows || !cols) return UH_OK;
    if (!C || (nnz && (!entries || !B))) return fail(UH_ERR_ARG, "uh_spmm_coo_f32: null buffer");
    if ((uintptr_t)entries & 3u) return fail(UH_ERR_ARG, "uh_spmm_coo_f32: entries must be 4-byte aligned");
    if (ldb < cols || ldc < cols) return fail(UH_ERR_ARG, "uh_spmm_coo_f32: leading dimension below the block");
    if (rows > 0x7fffffffu) return fail(UH_ERR_ARG, "uh_spmm_coo_f32: too many rows");
    int rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    UH_LAUNCH(k_spmm_coo, dim3(rows), dim3(256), 0, s, (const CooEntry*)entries, nnz, row0, B, ldb, C, ldc, cols);
    UH_LAUNCH_CHECK("k_spmm_coo");
    g_plain_launches++;
    return UH_OK;
}
