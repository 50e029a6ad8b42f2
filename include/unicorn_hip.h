/*
 * unicorn-mi355x — C-ABI of the HIP layer (libunicorn_hip.so).
 *
 * This is the only library that calls HIP. The pm:: runtime (pmStubHIP) and the apps' GPU
 * callbacks link to it; Python tests reach it through ctypes. It replaces, for one MI355X:
 *   - the reference's dlsym table over libcudart and its device helpers
 *     (source/code/src/pmCudaInterface.cu:69-378),
 *   - the generic H2D / launch / D2H sequence of one GPU subtask (pmCudaInterface.cu:474-550),
 *   - the CUDA/cuBLAS calls inside the apps' GPU callbacks
 *     (the testSuite apps' *CUDA.cu files, cited per launcher below).
 *
 * Conventions: plain pointers and sizes only; every function returns 0 on success and a
 * non-zero uh_status otherwise (uh_last_error() describes the last failure on the calling
 * thread). Streams are opaque `void*` (a hipStream_t). A stream created by
 * uh_batch_create() is a *batch handle*: kernel launchers called with it record a work item
 * in a device work queue instead of launching, and uh_batch_flush() runs all recorded items of
 * a family in one persistent launch. Launchers never synchronise, except uh_invert_index, which
 * waits for its stream to report malformed input.
 */
#ifndef UNICORN_HIP_H
#define UNICORN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum uh_status {
    UH_OK = 0,
    UH_ERR_HIP = 1,          /* a HIP runtime call failed */
    UH_ERR_ARG = 2,          /* invalid argument (shape, stride, pointer) */
    UH_ERR_OOM = 3,          /* device or pinned allocation failed */
    UH_ERR_NODEV = 4,        /* no HIP device */
    UH_ERR_BATCH = 5         /* operation not supported on a batch handle */
};

const char* uh_last_error(void);

/* ---- device (reference pmCudaInterface.cu:227-334: probe, bind, memory queries) ---- */
int uh_device_count(int* count);
int uh_set_device(int device);
int uh_device_props(int device, char* name, size_t name_len, size_t* total_mem, int* cu_count,
                    int* clock_khz, int* l2_bytes, int* lds_per_block);
int uh_mem_info(size_t* free_bytes, size_t* total_bytes);
int uh_device_sync(void);

/* ---- memory ---- */
int uh_malloc(void** ptr, size_t bytes);
int uh_free(void* ptr);
int uh_host_alloc(void** ptr, size_t bytes);          /* pinned host memory (hipHostMalloc) */
int uh_host_free(void* ptr);
int uh_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int uh_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int uh_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* Strided 2-D copy; kind 0 = H2D, 1 = D2H, 2 = D2D. */
int uh_memcpy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                 size_t rows, int kind, void* stream);
/* Fill n 32-bit words with a value using a kernel (never a memset node: graph-safe). */
int uh_fill_u32(void* dst, uint32_t value, size_t n_words, void* stream);

/* ---- streams, events, graphs ---- */
int uh_stream_create(void** stream);
int uh_stream_destroy(void* stream);
int uh_stream_sync(void* stream);
int uh_stream_query(void* stream);                    /* 0 = idle, 1 = busy */
int uh_stream_wait_event(void* stream, void* event);
int uh_event_create(void** event);
int uh_event_destroy(void* event);
int uh_event_record(void* event, void* stream);
int uh_event_sync(void* event);
int uh_event_query(void* event);                      /* 0 = complete, 1 = pending */
int uh_event_elapsed_ms(float* ms, void* start, void* stop);
int uh_graph_begin_capture(void* stream);
int uh_graph_end_capture(void* stream, void** graph_exec);
int uh_graph_launch(void* graph_exec, void* stream);
int uh_graph_destroy(void* graph_exec);

/* Launch a __global__ kernel of the pm kernel-kind callback type
 * (reference pmCudaInterface.cu:532-543). args = array of pointers to each argument. */
int uh_launch_kernel(const void* func, int gx, int gy, int gz, int bx, int by, int bz,
                     size_t shared_mem, void* stream, void** args);

/* ---- matrixTranspose (reference matrixTranspose_cuda, matrixTransposeCUDA.cu:46-69) ----
 * out[c * ld_out + r] = in[r * ld_in + c] for a rows x cols block of 32-bit elements (leading
 * dimensions in elements; in and out must not alias). */
int uh_transpose32(const void* in, uint64_t ld_in, void* out, uint64_t ld_out, uint32_t rows, uint32_t cols,
                   void* stream);

/* ---- prefixSum (reference prefixSum_cuda / elemAdd_cuda, prefixSumCUDA.cu:68-165) ----
 * uh_scan_i32: out[i] = in[0] + ... + in[i] modulo 2^32, n <= 2^24 (in may equal out).
 * uh_add_scalar_i32: data[i] += *value modulo 2^32, value a device pointer. */
int uh_scan_i32(const int32_t* in, int32_t* out, uint64_t n, void* stream);
int uh_add_scalar_i32(int32_t* data, uint64_t n, const int32_t* value, void* stream);

/* ---- sparseSolver (reference cusparseXcoo2csr + csrmm2, sparseSolverCUDA.cu:206-290) ----
 * entries: nnz records {int32 row, int32 col, float value}, sorted by row. C[r][j] = sum over the
 * entries k of row row0 + r (in order) of value[k] * B[col[k]][j], for r < rows, j < cols.
 * Products and sums are rounded separately (no fused multiply-add). */
int uh_spmm_coo_f32(const void* entries, uint32_t nnz, int32_t row0, uint32_t rows, const float* B, uint64_t ldb,
                    float* C, uint64_t ldc, uint32_t cols, void* stream);

/* ---- fft (reference cufftPlan1d / cufftPlanMany + cufftExecC2C, fftCUDA.cu:112-200) ----
 * Forward, unnormalised DFTs X[k] = sum_j x[j] exp(-2 pi i j k / n) of `count` sequences of n
 * single-precision complex points (float pairs), n a power of two in [2, 8192]. Element j of
 * sequence s is at complex index s * seq_stride + j * elem_stride of its buffer. in may equal out
 * when both stride pairs are equal; otherwise the buffers must not overlap. */
int uh_fft_c2c_f32(const void* in, uint64_t in_seq_stride, uint64_t in_elem_stride, void* out,
                   uint64_t out_seq_stride, uint64_t out_elem_stride, uint32_t n, uint32_t count, void* stream);

/* ---- pageRankInvertedIndex (the reference's GPU callbacks invertIndex_cudaLaunchFunc /
 *      pageRank_cudaLaunchFunc are empty, pageRankInvertedIndexCUDA.cu:92-100; these do what its
 *      CPU callbacks invertIndex_cpu / pageRank_cpu do, pageRankInvertedIndex.cpp:301-382, 438-472) ----
 * uh_invert_index: `pages` web-dump records (uint32: out-link count c <= max_out, then c 1-based
 * targets, record length max_out + 1) of the pages first_page.. become 12-byte entries
 * {src page, dest page (0-based), src out-links} in out_entries, grouped by dest / pages_per_subtask
 * in ascending group order, each group in reverse enumeration order (page, then link); totals[g]
 * receives the entry count of group g (groups <= 8192). Device pointers; the call waits for its
 * stream and fails if a record or link was malformed (those are skipped).
 * uh_prii_rank: out[0 .. out_pages) = 0, then out[dest - first_page] += damping * (x / out-links)
 * for each entry, x = in_ranks[src] (or `initial` when in_ranks is NULL); entries outside the
 * range are skipped. Float atomics: summation order is not fixed. */
int uh_invert_index(const uint32_t* dump, uint32_t pages, uint32_t max_out, uint32_t first_page, uint32_t total_pages,
                    uint32_t pages_per_subtask, uint32_t groups, void* out_entries, uint32_t* totals, void* stream);
int uh_prii_rank(const void* entries, uint32_t n, const float* in_ranks, float initial, float damping,
                 uint32_t first_page, float* out, uint32_t out_pages, uint32_t total_pages, void* stream);

/* ---- sentinel compression (replaces pmCudaInterface::CompressForSentinel,
 *      pmCudaInterface.cu:380-472, called from pmExecutionStub.cpp:3850-3881) ----
 * Counts the non-zero elements of data[0 .. elems) (type: pmReductionDataType, 0 int .. 5 double;
 * an element is zero when it compares equal to zero of its type) into *nonzero, waiting for the
 * stream. When *nonzero <= max_nonzero, also writes, in stream order, the compressed form to out
 * (8-byte aligned, at least 8 + max_nonzero * (element size + 4) bytes): a uint64 count, the
 * non-zero elements in index order, then their uint32 indices; otherwise out is not written.
 * elems must be below 2^32. */
int uh_sentinel_compress(const void* data, uint64_t elems, int type, void* out, uint64_t max_nonzero,
                         uint64_t* nonzero, void* stream);

/* ---- subscription staging (replaces the pinned-buffer host memcpy + per-record H2D of
 *      pmExecutionStub.cpp:3655-3847 for HBM-resident address spaces) ----
 * Copies n_records byte ranges: record i moves len[i] bytes from src + src_off[i] to
 * dst + dst_off[i]. The three arrays are device pointers to uint64. */
int uh_copy_records(void* dst, const void* src, const uint64_t* src_off, const uint64_t* dst_off,
                    const uint64_t* len, uint32_t n_records, uint64_t max_len, void* stream);

/* ---- built-in reductions (reference pmReducer::ReduceMemories, pmReducer.cpp:593-703) ----
 * dst[i] = dst[i] (op) src[i] for i < count elements; op is a pmReductionOpType value
 * (REDUCE_ADD .. REDUCE_BITWISE_XOR), type one of the UH_REDUCE_* element types below (in the
 * order of pmReductionDataType). Both buffers are device memory, 16-byte aligned. */
enum { UH_REDUCE_INT32 = 0, UH_REDUCE_UINT32 = 1, UH_REDUCE_INT64 = 2, UH_REDUCE_UINT64 = 3,
       UH_REDUCE_FLOAT32 = 4, UH_REDUCE_FLOAT64 = 5 };
int uh_reduce(void* dst, const void* src, size_t count, int op, int type, void* stream);
/* REDUCE_ADD over floats (uh_reduce with op 0, UH_REDUCE_FLOAT32) */
int uh_reduce_add_f32(float* dst, const float* src, size_t n, void* stream);

/* ---- imageFiltering (reference imageFilteringCUDA.cu:86-342) ----
 * One subtask: output tile rows [row0, row0+rows) x cols [col0, col0+cols) of an H x W RGB
 * image. The input view holds the halo'd region in inverted (bottom-up) row order: view row v
 * is image row (view_last_row - v), view column u is image column (view_first_col + u).
 * out(i,j,c) = sum_{dy,dx in [-r,r]} w[dy+r][dx+r] * in(clamp(i+dy), clamp(j+dx), 2-c)
 * mode 0 (REF): accumulated modulo 256, as the reference's char accumulators.
 * mode 1 (CFG): int32 sum, arithmetic right shift by `shift`, clamped to [0,255].
 * weights: (2r+1)^2 int8 values, row-major, read from host memory at call time. */
typedef struct uh_imagefilter_args {
    const uint8_t* in;  uint64_t in_stride;     /* bytes between view rows */
    uint8_t* out;       uint64_t out_stride;    /* bytes between output rows */
    int32_t row0, col0, rows, cols;             /* output tile in image coordinates */
    int32_t view_last_row, view_first_col;      /* input view origin in image coordinates */
    int32_t height, width;                      /* image size */
    int32_t radius, mode, shift, pad;
    int8_t weights[32 * 32];                    /* row-major (2r+1)x(2r+1), r <= 15 */
} uh_imagefilter_args;
int uh_imagefilter(const uh_imagefilter_args* args, void* stream);
/* Engine of radius 1 and 2 for launches and work items recorded afterwards: 1 = int8 MFMA
 * (default), 0 = v_dot4 (UH_FILTER_ENGINE=dot4 at start); -1 only reports. Both give the same
 * bytes. *previous (if not NULL) receives the engine before the call. */
int uh_imagefilter_engine(int engine, int* previous);

/* ---- pageRank pull SpMV over an HBM-resident in-link CSR (replaces
 *      pageRankCUDA.cu:46-164: push with float atomics over a padded web dump) ----
 * For destination pages t in [row0, row0+rows):
 *   y[t] = sum_{e in [row_ptr[t], row_ptr[t+1])} w[col[e]]
 *   w_next[t] = (deg[t] > 0) ? (damping * y[t]) / deg[t] : 0
 * where w[src] = (damping * x[src]) / deg[src] was produced by the previous iteration
 * (the reference's increment, pageRank.cpp:262,367). y and w_next are indexed by global page id.
 *
 * Source-blocked layout (bcnt != NULL). Pages form row groups of UH_PR_GROUP_ROWS destinations
 * and UH_PR_BLOCKS equal source blocks (block size rounded up to 32 pages). Within a row group
 * the in-links are ordered by (source block, destination, source):
 *   bseg[g * UH_PR_BLOCKS + x]  first link of group g, block x (bseg[groups * UH_PR_BLOCKS] = E);
 *   bcnt[(g * UH_PR_BLOCKS + x) * UH_PR_GROUP_ROWS + r]  links of row g*GROUP_ROWS + r in block x
 *   (u8; rows past the last page count 0; bcnt 4-byte aligned).
 * col is indexed as before (global link id) and row_ptr is not read. Each row's sum is taken per
 * block in link order and the blocks are added in block order, so results are deterministic.
 * With bcnt == NULL col is a plain in-link CSR (sources ascending within each row). */
#define UH_PR_GROUP_ROWS 2048
#define UH_PR_BLOCKS 8
typedef struct uh_pagerank_args {
    const uint32_t* row_ptr; const uint32_t* col; const uint32_t* deg;
    const float* w; float* y; float* w_next;
    uint32_t row0, rows;
    float damping; uint32_t pad;
    const uint8_t* bcnt; const uint32_t* bseg;   /* source-blocked layout, or NULL */
} uh_pagerank_args;
int uh_pagerank_pull(const uh_pagerank_args* args, void* stream);
/* Allocation-free variant of uh_pagerank_pull. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results as uh_pagerank_pull on a stream
 * (a batch handle is rejected with UH_ERR_BATCH); it calls no hipMalloc/hipFree. The
 * source-blocked layout's partial sums live in a static device array of 32 MiB per GPU (held until
 * the process ends), which limits one call to 512 row groups of UH_PR_GROUP_ROWS (rows
 * [row0, row0 + rows) may touch at most 512 groups: 1,048,576 rows when row0 is group-aligned);
 * larger calls return UH_ERR_ARG. Every call on one GPU shares that array: calls that overlap in
 * time on one GPU (on different streams) race. The CSR layout uses no array. */
int uh_pagerank_pull_noalloc(const uh_pagerank_args* args, void* stream);
/* w[i] = (damping * x0) / deg[i] (or 0 for dangling pages) for the first iteration. */
int uh_pagerank_init_w(float* w, const uint32_t* deg, uint32_t n, float x0, float damping,
                       void* stream);

/* ---- pageRank over the binned (propagation-blocked) layout (replaces pageRankCUDA.cu:46-164) ----
 * Pages form units of at most UH_PRB_UNIT_MAX consecutive pages (unit_start, K+1 entries); a unit
 * is a destination bin and a source chunk. Link values (the contribution w[src] of each link)
 * are stored bin-major: bin b's links occupy [bin_start[b], bin_start[b+1]) of values and dst16
 * (bin_start entries are multiples of 4; padding slots hold value 0, destination 0), grouped by
 * source chunk. A run is the links of one chunk into one bin: the runs of chunk c are
 * [run_ptr[c], run_ptr[c+1]); run r takes its sources from src16[run_off[r] .. run_off[r+1]) and
 * its values go to values[run_pos[r] ..]. dst16 = destination - unit_start[b], src16 = source -
 * unit_start[c]. For units [unit0, unit0 + units):
 *   kind 0 (first iteration): y = x0;  kind 1: y[t] = sum of the values of bin t's links, added
 *   as floor(v * 2^scale_exp) in 64-bit integers and converted back (order-independent);
 *   w[t] = deg[t] ? (damping * y[t]) / deg[t] : 0 for the unit's pages, and, if values_out is not
 *   NULL, values_out[run_pos[r] + k] = w[unit_start[c] + src16[run_off[r] + k]] for the unit's runs.
 * scale_exp must keep every row sum below 2^62 (damping <= 1: scale_exp <= 62 - ceil(log2(N x0))).
 * Pointers are indexed globally (page id, link slot, run id); values_in must be 16-byte and dst16
 * 8-byte aligned at index 0. values_in / dst16 (run_pos / src16) are read only for the units'
 * slots (runs) and may be NULL when there are none. */
#define UH_PRB_UNIT_MAX 16384
typedef struct uh_pagerank_binned_args {
    const uint32_t* unit_start; const uint32_t* bin_start;
    const uint32_t* run_ptr; const uint32_t* run_off; const uint32_t* run_pos;
    const uint16_t* dst16; const uint16_t* src16;
    const uint32_t* deg;
    const float* values_in; float* values_out; float* y;
    uint32_t unit0, units;
    float damping, x0;
    int32_t kind, scale_exp;
} uh_pagerank_binned_args;
int uh_pagerank_binned(const uh_pagerank_binned_args* args, void* stream);
/* Push formulation, as the reference callback: y[t-1] += (damping*x[src])/deg over a padded
 * record slice (count, t1..t_count, 0-pad to max_out) for pages [0, pages); float atomics. */
int uh_pagerank_push_ell(const uint32_t* records, uint32_t pages, uint32_t max_out,
                         const float* x, float x0, int first_iteration, float damping,
                         float* y, void* stream);

/* ---- dense GEMM, row-major: C = alpha * A(MxK) * B(KxN) + beta * C ----
 * precision 0 = split-BF16 3-term MFMA (hi*hi + hi*lo + lo*hi, fp32 accumulate),
 * precision 1 = FP32 MFMA (v_mfma_f32_*_f32), precision 2 = FP32 VALU reference kernel.
 * Column-major callers pass the transposed problem (C^T = B^T A^T).
 * Replaces cublasSgemm in matrixMultiplyBlasCUDA.cu:54-77 and luDecompositionCUDA.cu:190. */
typedef struct uh_gemm_args {
    const float* A; const float* B; float* C;
    uint64_t lda, ldb, ldc;
    int32_t M, N, K, precision;
    float alpha, beta;
} uh_gemm_args;
int uh_gemm(const uh_gemm_args* args, void* stream);
/* Tile configuration for a shape class (log2 buckets of M, N, K) and precision. */
int uh_gemm_set_tile(int M, int N, int K, int precision, int tile_id);
int uh_gemm_tile_count(int* n);
/* Autotuner (no reference counterpart; the reference calls cuBLAS): times every tile
 * configuration on an M x N x K product in scratch device memory (`reps` launches each, after one
 * warm-up) on a private stream, records the fastest for the shape class with uh_gemm_set_tile,
 * and returns it (and its time) when the pointers are non-NULL. Synchronous. A shape class that
 * was already tuned is not timed again unless `force`. */
int uh_gemm_autotune(int M, int N, int K, int precision, int reps, int force, int* tile_id, float* ms);

/* ---- LU (column-major, no pivoting, reference convention: L lower with the pivots on its
 *      diagonal, U unit upper; luDecomposition.cpp:116-125) ---- */
/* Factor one n x n diagonal block in place (reference luDecompositionCUDA.cu:66-97). */
int uh_lu_panel(float* a, uint64_t lda, int n, void* stream);
/* X := L^-1 X, L lower non-unit (n x n), X n x m (reference luDecompositionCUDA.cu:127). */
int uh_trsm_lln(const float* l, uint64_t ldl, float* x, uint64_t ldx, int n, int m, void* stream);
/* X := X U^-1, U upper unit (n x n), X m x n (reference luDecompositionCUDA.cu:147). */
int uh_trsm_run(const float* u, uint64_t ldu, float* x, uint64_t ldx, int m, int n, void* stream);

/* ---- iterative refinement helpers (FP64 residual, FP32 triangular solves) ---- */
/* r = b - A x in FP64; A column-major n x n (fp32), x (fp64) and b (fp64) n x nrhs
 * column-major; r written as fp64 and also as fp32 (rf) for the correction solve. */
int uh_residual_f64(const float* a, uint64_t lda, int n, const double* x, const double* b,
                    int nrhs, double* r, float* rf, void* stream);
/* Solve (L U) z = rhs in place for nrhs columns (fp32), with the factors of uh_lu_panel. */
int uh_lu_solve_f32(const float* lu, uint64_t ld, int n, float* rhs, int nrhs, void* stream);
/* Start of a refinement: x = 0, r = b, rf = (float) b (n x nrhs each, column-major). */
int uh_refine_init(const double* b, double* x, double* r, float* rf, int n, int nrhs, void* stream);
/* x += z (z fp32, x fp64), and norms: out[0..nrhs) = max|r|, out[nrhs..2nrhs) = max|x|. */
int uh_refine_update(double* x, const float* z, int n, int nrhs, void* stream);
int uh_colmax_abs_f64(const double* v, int n, int ncols, double* out, void* stream);

/* ---- batching: device work queue + persistent kernels ---- */
int uh_batch_create(void** batch_handle, void* stream);
int uh_batch_destroy(void* batch_handle);
int uh_is_batch(void* stream);
/* Number of work items recorded since the last flush. */
int uh_batch_pending(void* batch_handle, int* n);
/* Upload the recorded work items and launch one persistent kernel per kernel family on the
 * batch's stream. With use_graph != 0 the upload+launch sequence is captured once per batch
 * shape and replayed as a hipGraph. */
int uh_batch_flush(void* batch_handle, int use_graph);
/* Device time of flushes: with every = N >= 1, events are recorded on the batch stream around the
 * persistent kernels of every N-th flush (the first one after the call, then each N-th; the work
 * queue's upload precedes the first event); every = 0 records none. uh_batch_timing waits for
 * the pending ones and returns the accumulated milliseconds and the number of flushes timed. */
int uh_batch_set_timing(void* batch_handle, int every);
int uh_batch_timing(void* batch_handle, double* total_ms, uint64_t* flushes, int reset);
/* The timed flushes that launched one kernel family only: their milliseconds, count and
 * algorithmic work (UH_FAMILY_GEMM: flop issued, 2 M N K per product, 3x for split-BF16;
 * UH_FAMILY_FILTER: bytes, 2 * rows * cols * 3 per item; other families 0). reset clears that
 * family's totals; uh_batch_timing's reset clears every family's. */
enum { UH_FAMILY_FILTER = 0, UH_FAMILY_GEMM = 1, UH_FAMILY_PAGERANK = 2, UH_FAMILY_TRSM = 3,
       UH_FAMILY_PAGERANK_BLOCKED = 4, UH_FAMILY_PAGERANK_BINNED = 5 };
int uh_batch_timing_family(void* batch_handle, int family, double* ms, uint64_t* flushes, double* work,
                           int reset);
/* Launch counters (kernels launched through this library since load). */
int uh_launch_counters(uint64_t* plain_launches, uint64_t* persistent_launches, uint64_t* items);

#ifdef __cplusplus
}
#endif

#endif
