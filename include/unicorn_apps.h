/*
 * unicorn-mi355x — C-ABI drivers of the four benchmark apps (libunicorn_apps.so).
 *
 * The apps themselves are written against the pm:: API (pmPublicDefinitions.h) exactly as the
 * reference's testSuite apps are: data-distribution callbacks subscribe memory, CPU and GPU
 * subtask callbacks compute. These drivers replace the reference's per-app main()/commonStart
 * harness (testSuite/common/source/code/src/commonAPI.cpp:142-298) with plain C entry points
 * that tests and benchmarks call through ctypes: create the address spaces, load inputs, submit
 * the task(s) on a chosen device set, fetch outputs.
 *
 * All functions return 0 on success; on failure a non-zero pmStatus (or -1 for a driver-level
 * error) and ua_last_error() describes it. Handles are opaque.
 */
#ifndef UNICORN_APPS_H
#define UNICORN_APPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* device sets, as the reference harness's local CPU / GPU / CPU+GPU modes */
enum { UA_DEVICE_CPU = 0, UA_DEVICE_GPU = 1, UA_DEVICE_BOTH = 2 };

typedef struct ua_stats {
    uint64_t subtasks, batches, h2d_bytes, d2h_bytes, d2d_bytes, cache_hits, evictions;
    uint64_t plain_launches, persistent_launches, queue_items;
    uint64_t ns_queue, ns_stage, ns_flush, ns_wait, ns_retire, ns_wake;   /* see pmHipStubStats */
    uint64_t ma_copies, ma_wins, oom_waits, sentinel_compressed;
} ua_stats;

int ua_init(int cpu_stubs);                   /* pmInitialize; cpu_stubs 0 = one per core */
int ua_finalize(void);
int ua_has_gpu(void);
int ua_set_hip_mode(int mode);                /* PM_HIP_MODE_PLAIN / RESIDENT / BATCHED */
int ua_set_hip_compact(int compact);          /* force COMPACT views in RESIDENT/BATCHED */
int ua_set_hip_graphs(int on);
int ua_set_hip_depth(int depth);
int ua_set_batch_limit(int subtasks);
int ua_set_hbm_budget(uint64_t bytes);
int ua_set_sentinel_threshold(uint64_t bytes, uint64_t* previous);   /* pmSetSentinelCompressionThreshold */
/* canSplitCpuSubtasks / canSplitGpuSubtasks of the imageFiltering and matrixMultiply tasks
 * submitted afterwards (default 1 / 0, as the reference apps) and the pieces per split subtask
 * (pmSetSplitFactors; 0 = one per stub of the splitting device type). */
int ua_set_split(int cpu, int gpu, int cpu_pieces, int gpu_pieces);
/* multiAssignEnabled of the imageFiltering and matrixMultiply tasks submitted afterwards
 * (default 1, as the reference). */
int ua_set_multi_assign(int on);
int ua_get_stats(ua_stats* s);
int ua_reset_stats(void);
int ua_set_kernel_timing(int every);           /* events around every every-th BATCHED flush (0: off) */
int ua_kernel_time(double* ms, uint64_t* flushes, int reset);
/* flushes that launched one kernel family only (UH_FAMILY_*): time, count, algorithmic work */
int ua_kernel_time_family(int family, double* ms, uint64_t* flushes, double* work, int reset);
int ua_device_sync(void);
const char* ua_last_error(void);

/* ---- imageFiltering (reference testSuite/imageFiltering) ----
 * Input: width*height*3 bytes, rows stored bottom-up (the reference's inverted layout,
 * imageFiltering.cpp:98-134); output: width*height*3 bytes top-down, channels swapped. */
typedef struct ua_if_config {
    int32_t width, height, radius;
    int32_t mode;          /* 0 = reference char arithmetic (mod 256), 1 = shift + clamp */
    int32_t shift;         /* right shift of mode 1 */
    int32_t tile;          /* subtask tile side (reference TILE_DIM 2048) */
    int8_t weights[32 * 32];   /* (2r+1)^2 values, row-major */
} ua_if_config;
void* ua_if_create(const ua_if_config* cfg);
int ua_if_set_input(void* h, const uint8_t* image);
int ua_if_run(void* h, int device, double* seconds);
int ua_if_get_output(void* h, uint8_t* out);
void* ua_if_clone(void* h);                  /* same input image, own output */
int ua_if_fetch(void* h, uint64_t* checksum); /* output to the host; checksum may be NULL */
int ua_if_get_input(void* h, uint8_t* image); /* the input image (rows bottom-up) */
int ua_if_get_config(void* h, ua_if_config* cfg);
int ua_if_destroy(void* h);

/* ---- matrixMultiply (reference testSuite/matrixMultiplyBlas decomposition, plain CPU method) ----
 * n x n row-major float32; subtasks = (n/block)^2 output blocks; precision as uh_gemm. */
void* ua_mm_create(int n, int block, int precision);
int ua_mm_set_inputs(void* h, const float* A, const float* B);
int ua_mm_run(void* h, int device, double* seconds);
int ua_mm_get_output(void* h, float* C);
void* ua_mm_clone(void* h);                  /* same A and B, own C */
int ua_mm_fetch(void* h, uint64_t* checksum);
int ua_mm_get_inputs(void* h, float* A, float* B);
int ua_mm_destroy(void* h);

/* ---- matrixTranspose (reference testSuite/matrixTranspose, square blocks of
 *      min(rows, cols, 2048)): int32 rows x cols, out of place (COMPACT views) or in place (one
 *      READ_WRITE space) ---- */
void* ua_mt_create(uint64_t rows, uint64_t cols, int inplace);
int ua_mt_set_input(void* h, const int32_t* data);   /* rows x cols, row-major */
int ua_mt_run(void* h, int device, double* seconds);
int ua_mt_get_output(void* h, int32_t* out);          /* cols x rows, row-major */
int ua_mt_destroy(void* h);

/* ---- prefixSum (reference testSuite/prefixSum): inclusive int32 prefix sum (wrap-around) by a
 *      chunk-scan task, a host scan of the chunk totals and an add task ---- */
void* ua_ps_create(uint32_t n, uint32_t elems_per_subtask);   /* reference: 1,000,000 per subtask */
int ua_ps_set_input(void* h, const int32_t* data);
int ua_ps_run(void* h, int device, double* seconds);
int ua_ps_get_output(void* h, int32_t* out);
int ua_ps_destroy(void* h);

/* ---- sparseSolver (reference testSuite/sparseSolver): C = A B, A sparse n x n in COO order
 *      sorted by (row, col), B dense, one subtask per square block of min(n, 2048) ---- */
void* ua_ss_create(uint64_t n, uint64_t nnz, const int32_t* rows, const int32_t* cols, const float* vals,
                   const float* B);
int ua_ss_run(void* h, int device, double* seconds);
int ua_ss_get_output(void* h, float* C);
int ua_ss_destroy(void* h);

/* ---- fft (reference testSuite/fft, 2-D without transposes): forward unnormalised 2-D DFT of a
 *      rows x cols complex float matrix (interleaved float pairs) by a row task and a column task
 *      of rows_per_subtask rows / columns per subtask (0: the reference's 512); rows and cols are
 *      powers of two in [2, 8192] ---- */
void* ua_fft_create(uint32_t rows, uint32_t cols, uint32_t rows_per_subtask);
int ua_fft_set_input(void* h, const float* data);
int ua_fft_run(void* h, int device, double* seconds);
int ua_fft_get_output(void* h, float* out);
int ua_fft_destroy(void* h);

/* ---- pageRankInvertedIndex (reference testSuite/pageRankInvertedIndex): pageRank over a web
 *      dump directory (the ua_webdump_* format) by an index-inversion task with data
 *      redistribution, then `iterations` (0: 2) pageRank tasks over the inverted index;
 *      pages_per_subtask 0 = the reference's 10,000,000, rounded to whole files ---- */
void* ua_prii_create(const char* dir, uint32_t pages_per_subtask);
int ua_prii_run(void* h, int device, uint32_t iterations, double* seconds);
/* the redistributed index: n_entries entries {src, dest, src out-links} (uint32 x 3) and the
 * inlink count of each of the subtasks (entries NULL: counts only) */
int ua_prii_get_index(void* h, uint32_t* entries, uint64_t capacity, uint64_t* n_entries, uint32_t* counts);
int ua_prii_get_ranks(void* h, float* ranks);
int ua_prii_info(void* h, uint32_t* pages, uint32_t* pages_per_subtask, uint32_t* subtasks);
int ua_prii_destroy(void* h);

/* ---- web graph generator (the role of the reference's webGenerator.pl) ----
 * Out-degree 0 with probability `dangling`, else P(k) ~ (k + k0)^-2 on [1, max_out] with k0
 * chosen for an overall mean of `mean_out`; targets by a Polya urn (in-degree + 1).
 * Call once with col == NULL to get row_ptr and *edges, then again with col of that size. */
int ua_webgen(uint32_t pages, uint64_t seed, double mean_out, uint32_t max_out, double dangling,
              uint32_t* row_ptr, uint32_t* col, uint64_t col_capacity, uint64_t* edges);
double ua_webgen_shift(double mean_out, uint32_t max_out, double dangling);

/* ---- pageRank (reference testSuite/pageRank): graph as an out-link CSR (0-based targets);
 * subtasks own `pages_per_subtask` destination pages; ua_pr_run runs the init task and
 * `iterations` iteration tasks; ranks are those after the last iteration. */
void* ua_pr_create(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                   uint32_t pages_per_subtask, float damping);
int ua_pr_run(void* h, int iterations, int device, double* seconds);
int ua_pr_get_ranks(void* h, float* ranks);
uint64_t ua_pr_edges(void* h);
void* ua_pr_clone(void* h);                  /* same graph, own rank state */
int ua_pr_fetch(void* h, uint64_t* checksum);
int ua_pr_destroy(void* h);
/* Graph address space (see uh_pagerank_args for the source-blocked layout). ua_pr_block_counts
 * fills bcnt (ceil(pages / UH_PR_GROUP_ROWS) * UH_PR_BLOCKS * UH_PR_GROUP_ROWS bytes) and returns
 * 1, or 0 when a page has more than 255 in-links from one source block or pages >= 2^30 (the
 * blocked kernels address the rank vector with 32-bit byte offsets). ua_pr_build_graph writes
 * [in_row_ptr | outdeg | in_col (| bseg | bcnt)] into ua_pr_graph_bytes(pages, edges, bcnt != 0)
 * bytes; in_col is source-blocked when bcnt is given. ua_pr_set_layout chooses the layout of later
 * ua_pr_create calls: 2 = binned (uh_pagerank_binned_args; the default, used when damping <= 1
 * and the link slots fit 32 bits), 1 = source-blocked, 0 = plain in-link CSR; a layout that cannot
 * hold a graph falls back to the next (binned -> blocked -> CSR). ua_pr_layout reports an instance's
 * layout, ua_pr_blocked whether it is the source-blocked one. ua_prb_layout plans (graph == NULL:
 * returns the graph bytes and info = {units, runs, slots, subtasks, offsets of the deg, unit_start,
 * bin_start, run_ptr, run_off, run_pos, dst16 and src16 regions}) or builds the binned graph
 * address space for `cus` CUs; it returns 0 when the layout cannot hold the graph. */
int ua_pr_block_counts(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col, uint8_t* bcnt);
uint64_t ua_pr_graph_bytes(uint32_t pages, uint64_t edges, int blocked);
int ua_pr_build_graph(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                      const uint8_t* bcnt, uint32_t* graph);
void ua_pr_set_layout(int layout);
int ua_pr_blocked(void* h);
int ua_pr_layout(void* h);
uint64_t ua_prb_layout(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                       uint32_t pages_per_subtask, uint32_t cus, void* graph, uint64_t* info);
/* The reference's on-disk web dump (pageRank.cpp:53-197, scripts/webGenerator.pl:121-274):
 * <dir>/web_dump_metadata = {pages, max out-links, pages per file} (uint32); <dir>/web/page_<n>
 * = fixed records {count, count 1-based targets, zero padding to max out-links}, file n holding
 * pages n, n+1, ... ua_webdump_read works like ua_webgen (col == NULL: row_ptr and *edges; then
 * col, 0-based); ua_webdump_write writes an out-link CSR in that format;
 * ua_pr_create_from_webdump builds a pageRank instance from a dump. */
int ua_webdump_meta(const char* dir, uint32_t* pages, uint32_t* max_out, uint32_t* per_file);
int ua_webdump_read(const char* dir, uint32_t* row_ptr, uint32_t* col, uint64_t col_capacity, uint64_t* edges);
int ua_webdump_write(const char* dir, uint32_t pages, uint32_t max_out, uint32_t per_file,
                     const uint32_t* row_ptr, const uint32_t* col);
void* ua_pr_create_from_webdump(const char* dir, uint32_t pages_per_subtask, float damping);
/* pageRank in the reference's decomposition (pageRank.cpp:237-378, 449-560): subtasks own
 * source pages and push into private full-length vectors folded by the built-in REDUCE_ADD
 * (in HBM on the HIP stub's resident modes). Records address space = the web dump's records;
 * ua_prp_create pads an out-link CSR to its largest out-degree, ua_prp_create_from_webdump copies
 * the dump's page files. ua_prp_run runs `iterations` iteration tasks from x0 = 1.0. */
void* ua_prp_create(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col,
                    uint32_t pages_per_subtask, float damping);
void* ua_prp_create_from_webdump(const char* dir, uint32_t pages_per_subtask, float damping);
int ua_prp_run(void* h, int iterations, int device, double* seconds);
int ua_prp_get_ranks(void* h, float* ranks);
int ua_prp_destroy(void* h);

/* ---- luDecomposition (reference testSuite/luDecomposition) + refinement solve ----
 * n x n float32 column-major, blocks of `block` (<= 512); precision 0 = split-BF16 trailing
 * update, 2 = fp32. ua_lu_solve: b, x n x nrhs float64 column-major (nrhs <= 8); info[0] = steps,
 * info[1] = final normwise backward error, info[2] = after the first step, info[3] = ||A||inf. */
void* ua_lu_create(int n, int block, int precision);
int ua_lu_set_matrix(void* h, const float* A);
int ua_lu_prefetch(void* h);                 /* matrix + pristine copy to HBM (no-op without GPU) */
int ua_lu_factor(void* h, int device, double* seconds);
int ua_lu_get_factors(void* h, float* LU);
int ua_lu_solve(void* h, const double* b, int nrhs, int max_steps, int device, double* x, double* info,
                double* seconds);
int ua_lu_destroy(void* h);

/* ---- serving (no reference counterpart: the request loop of the served mode) ----
 * A catalogue of datasets (ua_serve_add_* return a dataset id, or -1) and a stream of requests
 * (dataset id, arrival time in seconds from the start, sorted). `slots` worker threads serve the
 * stream in arrival order; a request runs one task of its dataset's app on a free clone (pageRank:
 * init + iterations), then, if `fetch`, brings its output to the host (and hashes it into
 * checksums[i] when checksums is not NULL). latency_ms[i] = completion - arrival. Returns 0 or
 * the first failing request's status. Outputs are FNV-1a-64 hashed 8 bytes at a time. */
void* ua_serve_create(void);
int ua_serve_add_mm(void* s, int n, int block, int precision, uint64_t seed);
int ua_serve_add_if(void* s, int side, int radius, int mode, int shift, int tile, uint64_t seed);
int ua_serve_add_pr(void* s, uint32_t pages, uint32_t pages_per_subtask, int iterations, uint64_t seed);
int ua_serve_run(void* s, int n, const int* dataset, const double* arrival_s, int slots, int device,
                 int fetch, double* latency_ms, uint64_t* checksums, double* wall_s);
int ua_serve_destroy(void* s);
/* The instance that owns dataset `ds`'s inputs (requests run on its clones, never on it), or NULL:
 * ua_mm_* / ua_if_* / ua_pr_* handle by the dataset's kind, for checking requests' outputs. */
void* ua_serve_dataset(void* s, int ds);
/* The output hash that ua_*_fetch and ua_serve_run report (FNV-1a-64 over 8-byte words, then
 * the tail bytes). */
uint64_t ua_hash(const void* p, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif
