/*
 * ORACLE — test infrastructure only.
 *
 * Plain-C restatement of the reference's CPU algorithms for the device subtask path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. Nothing in
 * the product (unicorn_amd/) links, imports or calls this code.
 *
 * Functions and the reference text each follows (paths relative to the reference repository):
 *   or_imagefilter        testSuite/imageFiltering/source/code/src/imageFiltering.cpp:137-211
 *                         (serialImageFilter) + the config's shift/clamp mode
 *   or_imagefilter_tiled  imageFiltering.cpp:279-472 (GetSubtaskSubscription, imageFilter_cpu)
 *                         over the tile decomposition, one pthread per core
 *   or_pagerank           testSuite/pageRank/source/code/src/pageRank.cpp:336-378 (pageRank_cpu)
 *                         per subtask + pmReducer.cpp:593-605 (REDUCE_ADD) in subtask order, and
 *                         pageRank.cpp:237-276 (serialPageRank) when subtask_pages == 0
 *   or_pagerank_f64       the same recurrence in double (FP64 oracle)
 *   or_matmul_blocks      testSuite/matrixMultiplyBlas/source/code/src/matrixMultiplyBlas.cpp:90-133
 *                         block decomposition with the plain i-j-k loop of
 *                         testSuite/matrixMultiply/source/code/src/matrixMultiply.cpp:47-55 in float
 *   or_matmul_f64         C = A B in double (FP64 oracle)
 *   or_lu                 testSuite/luDecomposition/source/code/src/luDecomposition.cpp:116-125
 *                         (serialLUDecomposition: sscal then sger, column-major)
 *   or_lu_f64             the same factorisation in double
 *   or_lu_solve_f64       forward/backward substitution with those factors (FP64)
 *   or_matvec_f64         r = b - A x in double
 *   or_transpose_i32      testSuite/matrixTranspose/source/code/src/matrixTranspose.cpp:187-261
 *                         (serialMatrixTranspose): the element at output index i of the
 *                         cols x rows result is input index (i * cols) mod (rows * cols - 1),
 *                         the map its cycle-following loop walks; the last element stays
 *   or_prefix_sum_i32     testSuite/prefixSum/source/code/src/prefixSum.cpp:47-53 (serialPrefixSum)
 *                         over int (wrap-around), and the chunked form of 225-290 (prefixSum_cpu
 *                         per chunk, PopulateAuxMem, elemAdd_cpu) when chunk > 0
 *   or_spmm_coo           testSuite/sparseSolver/source/code/src/sparseSolver.cpp:78-85, 189-263
 *                         (serialSparseMatrixMultiply: MKL coomm, C = A B, beta 0) restated as a
 *                         loop over the COO entries in order, in float (round each product and
 *                         sum), or in double when dbl != 0
 *   or_transpose_blocks   matrixTranspose.cpp:294-378 (data distribution + matrixTranspose_cpu,
 *                         square blocks): every bs x bs block of the input moves to its
 *                         transposed position, one subtask at a time
 *   or_fft2_f64           testSuite/fft/source/code/src/fft.cpp:115-155 (fftSerial, FFT_2D: FFTW
 *                         forward row transforms, transpose, row transforms, transpose), i.e. the
 *                         unnormalised 2-D DFT X[u][v] = sum x[r][c] exp(-2 pi i (r u/R + c v/C));
 *                         FFTW 3.3.3 (vendored under testSuite/fft/thirdparty, not built: it needs
 *                         its own configure step) is restated by its definition, in double, as
 *                         radix-2 transforms of the rows, then of the columns
 *   or_prii_invert        testSuite/pageRankInvertedIndex/source/code/src/pageRankInvertedIndex.cpp:
 *                         301-399 (invertIndex_cpu: per-group counts, each group filled from its
 *                         top; invertIndexDataRedistribution: group g -> order g) and the runtime's
 *                         redistribution (pmRedistributor: orders ascending, subtasks ascending)
 *   or_prii_pagerank      pageRankInvertedIndex.cpp:438-472 (pageRank_cpu) per subtask over the
 *                         redistributed index, PAGE_RANK_ITERATIONS tasks (InvertIndexTask /
 *                         ParallelPageRankIteration, 546-637)
 *   or_prii_serial        pageRankInvertedIndex.cpp:208-247 (serialPageRank)
 *
 * Parity pins (tests/test_oracle.py): golden vectors produced by the reference's own serial
 * functions compiled from its sources (oracle/Makefile, oracle/_ref), closed forms and brute force.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ image filter */

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* One output row range [r0, r1) of the whole-image filter. in: H x W x 3, rows inverted
 * (image row m at memory row H-1-m); out: top-down rows, channel c = sum of input channel 2-c.
 * mode 0: char accumulation (mod 256), as the reference. mode 1: int sum >> shift, clamp. */
static void filter_rows(const uint8_t* in, int W, int H, const int8_t* w, int r, int mode, int shift,
                        uint8_t* out, int r0, int r1) {
    const int D = 2 * r + 1;
    for (int i = r0; i < r1; ++i) {
        for (int j = 0; j < W; ++j) {
            int acc[3] = {0, 0, 0};
            signed char cacc[3] = {0, 0, 0};
            for (int k = i - r; k <= i + r; ++k) {
                for (int l = j - r; l <= j + r; ++l) {
                    const int m = clampi(k, 0, H - 1), n = clampi(l, 0, W - 1);
                    const size_t idx = ((size_t)(H - 1 - m) * W + n) * 3;
                    const int wt = w[(k - (i - r)) * D + (l - (j - r))];
                    for (int q = 0; q < 3; ++q) {
                        if (mode == 0) cacc[q] = (signed char)(cacc[q] + (signed char)in[idx + q] * wt);
                        else acc[q] += (int)in[idx + q] * wt;
                    }
                }
            }
            const size_t o = ((size_t)i * W + j) * 3;
            for (int q = 0; q < 3; ++q) {
                if (mode == 0) out[o + q] = (uint8_t)cacc[2 - q];
                else out[o + q] = (uint8_t)clampi(acc[2 - q] >> shift, 0, 255);
            }
        }
    }
}

void or_imagefilter(const uint8_t* in, int W, int H, const int8_t* w, int r, int mode, int shift, uint8_t* out) {
    filter_rows(in, W, H, w, r, mode, shift, out, 0, H);
}

struct filter_job {
    const uint8_t* in; int W, H; const int8_t* w; int r, mode, shift, tile; uint8_t* out;
    int ntiles; int* next; pthread_mutex_t* mtx;
};

/* Tile subtask as the reference CPU callback computes it: the halo'd input view and the output
 * tile, using clamp-to-edge against the image bounds (imageFiltering.cpp:419-469). */
static void filter_tile(const struct filter_job* J, int t) {
    const int tx = (J->W + J->tile - 1) / J->tile;
    const int sc = (t % tx) * J->tile, sr = (t / tx) * J->tile;
    const int ec = sc + J->tile < J->W ? sc + J->tile : J->W;
    const int er = sr + J->tile < J->H ? sr + J->tile : J->H;
    const int D = 2 * J->r + 1;
    for (int i = sr; i < er; ++i) {
        for (int j = sc; j < ec; ++j) {
            int acc[3] = {0, 0, 0};
            signed char cacc[3] = {0, 0, 0};
            for (int k = i - J->r; k <= i + J->r; ++k) {
                const int m = clampi(k, 0, J->H - 1);
                const uint8_t* row = J->in + (size_t)(J->H - 1 - m) * J->W * 3;
                for (int l = j - J->r; l <= j + J->r; ++l) {
                    const int n = clampi(l, 0, J->W - 1);
                    const int wt = J->w[(k - (i - J->r)) * D + (l - (j - J->r))];
                    const uint8_t* px = row + (size_t)n * 3;
                    for (int q = 0; q < 3; ++q) {
                        if (J->mode == 0) cacc[q] = (signed char)(cacc[q] + (signed char)px[q] * wt);
                        else acc[q] += (int)px[q] * wt;
                    }
                }
            }
            uint8_t* o = J->out + ((size_t)i * J->W + j) * 3;
            for (int q = 0; q < 3; ++q)
                o[q] = J->mode == 0 ? (uint8_t)cacc[2 - q] : (uint8_t)clampi(acc[2 - q] >> J->shift, 0, 255);
        }
    }
}

static void* filter_worker(void* arg) {
    struct filter_job* J = (struct filter_job*)arg;
    for (;;) {
        pthread_mutex_lock(J->mtx);
        int t = (*J->next)++;
        pthread_mutex_unlock(J->mtx);
        if (t >= J->ntiles) break;
        filter_tile(J, t);
    }
    return NULL;
}

/* Tiles [t0, t1) of the decomposition on `threads` threads (one subtask at a time each). */
void or_imagefilter_tiled(const uint8_t* in, int W, int H, const int8_t* w, int r, int mode, int shift,
                          int tile, uint8_t* out, int t0, int t1, int threads) {
    pthread_mutex_t mtx = PTHREAD_MUTEX_INITIALIZER;
    int next = t0;
    struct filter_job J = {in, W, H, w, r, mode, shift, tile, out, t1, &next, &mtx};
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, filter_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
}

/* ------------------------------------------------------------------ pageRank */

/* Graph in the reference's record semantics: page p has out_deg = row_ptr[p+1]-row_ptr[p]
 * targets col[row_ptr[p] ...] given as 0-based page ids (the reference stores them 1-based).
 * One iteration of subtask s over source pages [s*P, min((s+1)*P, N)):
 *   y_s[t] += (0.85f * x[p]) / (float)deg(p)    (x = 1.0 on iteration 0)
 * then y = ((y_0 + y_1) + y_2) + ...  (REDUCE_ADD in subtask order).
 * subtask_pages == 0: one pass over all pages (serialPageRank). */
void or_pagerank(uint32_t N, const uint32_t* row_ptr, const uint32_t* col, int iterations,
                 uint32_t subtask_pages, float damping, float* y_out) {
    float* x = (float*)malloc(sizeof(float) * (size_t)N);
    float* y = (float*)malloc(sizeof(float) * (size_t)N);
    float* part = (float*)malloc(sizeof(float) * (size_t)N);
    const uint32_t P = subtask_pages ? subtask_pages : N;
    for (int it = 0; it < iterations; ++it) {
        memset(y, 0, sizeof(float) * (size_t)N);
        for (uint32_t s0 = 0; s0 < N; s0 += P) {
            const uint32_t s1 = s0 + P < N ? s0 + P : N;
            float* dst = (s0 == 0) ? y : part;
            if (s0 != 0) memset(part, 0, sizeof(float) * (size_t)N);
            for (uint32_t p = s0; p < s1; ++p) {
                const uint32_t deg = row_ptr[p + 1] - row_ptr[p];
                const float incr = (float)(damping * (it == 0 ? 1.0f : x[p])) / (float)deg;
                for (uint32_t e = row_ptr[p]; e < row_ptr[p + 1]; ++e) dst[col[e]] += incr;
            }
            if (s0 != 0)
                for (uint32_t t = 0; t < N; ++t) y[t] += part[t];
        }
        float* tmp = x; x = y; y = tmp;
    }
    memcpy(y_out, x, sizeof(float) * (size_t)N);
    free(x); free(y); free(part);
}

void or_pagerank_f64(uint32_t N, const uint32_t* row_ptr, const uint32_t* col, int iterations, double damping,
                     double* y_out) {
    double* x = (double*)malloc(sizeof(double) * (size_t)N);
    double* y = (double*)malloc(sizeof(double) * (size_t)N);
    for (int it = 0; it < iterations; ++it) {
        memset(y, 0, sizeof(double) * (size_t)N);
        for (uint32_t p = 0; p < N; ++p) {
            const uint32_t deg = row_ptr[p + 1] - row_ptr[p];
            if (!deg) continue;
            const double incr = damping * (it == 0 ? 1.0 : x[p]) / (double)deg;
            for (uint32_t e = row_ptr[p]; e < row_ptr[p + 1]; ++e) y[col[e]] += incr;
        }
        double* tmp = x; x = y; y = tmp;
    }
    memcpy(y_out, x, sizeof(double) * (size_t)N);
    free(x); free(y);
}

/* ------------------------------------------------------------------ matrix multiply */

struct mm_job {
    const float* A; const float* B; float* C; int n, bs, nb; int* next; pthread_mutex_t* mtx; int s1;
};

/* Subtask s: blockRow = s % bpd, blockCol = s / bpd (matrixMultiplyBlas.cpp:94-97); the block is
 * computed with the plain i-j-k loop, C pre-zeroed, k ascending (matrixMultiply.cpp:47-55). */
static void mm_block(const struct mm_job* J, int s) {
    const int bpd = J->nb, br = s % bpd, bc = s / bpd, n = J->n, bs = J->bs;
    for (int i = br * bs; i < (br + 1) * bs; ++i)
        for (int j = bc * bs; j < (bc + 1) * bs; ++j) {
            float c = 0.f;
            for (int k = 0; k < n; ++k) c += J->A[(size_t)i * n + k] * J->B[(size_t)k * n + j];
            J->C[(size_t)i * n + j] = c;
        }
}

static void* mm_worker(void* arg) {
    struct mm_job* J = (struct mm_job*)arg;
    for (;;) {
        pthread_mutex_lock(J->mtx);
        int s = (*J->next)++;
        pthread_mutex_unlock(J->mtx);
        if (s >= J->s1) break;
        mm_block(J, s);
    }
    return NULL;
}

/* Subtasks [s0, s1) of the N x N product with block side bs, on `threads` threads. */
void or_matmul_blocks(const float* A, const float* B, float* C, int n, int bs, int s0, int s1, int threads) {
    pthread_mutex_t mtx = PTHREAD_MUTEX_INITIALIZER;
    int next = s0;
    struct mm_job J = {A, B, C, n, bs, n / bs, &next, &mtx, s1};
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, mm_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
}

/* General row-major C(MxN) = A(MxK) B(KxN) in double. */
void or_matmul_f64(const float* A, const float* B, double* C, int M, int N, int K) {
    for (int i = 0; i < M; ++i) {
        double* c = C + (size_t)i * N;
        for (int j = 0; j < N; ++j) c[j] = 0.0;
        for (int k = 0; k < K; ++k) {
            const double a = A[(size_t)i * K + k];
            const float* b = B + (size_t)k * N;
            for (int j = 0; j < N; ++j) c[j] += a * (double)b[j];
        }
    }
}

/* ------------------------------------------------------------------ LU */

/* In-place LU of an n x n column-major block with leading dimension ld: for each i,
 * row i right of the diagonal is scaled by (float)(1.0 / a_ii) (sscal with a double reciprocal
 * converted to float), then a_rc -= a_ri * a_ic for r, c > i (sger, alpha = -1: multiply and
 * add, no fused multiply-add). */
void or_lu(float* a, int n, int ld) {
    for (int i = 0; i < n - 1; ++i) {
        const float d = (float)(1.0 / (double)a[i + (size_t)i * ld]);
        for (int c = i + 1; c < n; ++c) a[i + (size_t)c * ld] *= d;
        for (int c = i + 1; c < n; ++c) {
            const float u = -a[i + (size_t)c * ld];
            float* col = a + (size_t)c * ld;
            const float* li = a + (size_t)i * ld;
            for (int r = i + 1; r < n; ++r) {
                volatile float p = li[r] * u;
                col[r] = col[r] + p;
            }
        }
    }
}

void or_lu_f64(double* a, int n, int ld) {
    for (int i = 0; i < n - 1; ++i) {
        const double d = 1.0 / a[i + (size_t)i * ld];
        for (int c = i + 1; c < n; ++c) a[i + (size_t)c * ld] *= d;
        for (int c = i + 1; c < n; ++c) {
            const double u = a[i + (size_t)c * ld];
            double* col = a + (size_t)c * ld;
            const double* li = a + (size_t)i * ld;
            for (int r = i + 1; r < n; ++r) col[r] -= li[r] * u;
        }
    }
}

/* Solve (L U) x = b with factors stored as above (L lower incl. diagonal, U unit upper);
 * nrhs column-major right-hand sides of length n. */
void or_lu_solve_f64(const double* lu, int n, int ld, double* x, int nrhs) {
    for (int k = 0; k < nrhs; ++k) {
        double* v = x + (size_t)k * n;
        for (int i = 0; i < n; ++i) {
            double s = v[i];
            for (int j = 0; j < i; ++j) s -= lu[i + (size_t)j * ld] * v[j];
            v[i] = s / lu[i + (size_t)i * ld];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = v[i];
            for (int j = i + 1; j < n; ++j) s -= lu[i + (size_t)j * ld] * v[j];
            v[i] = s;
        }
    }
}

/* r = b - A x, A fp32 column-major, all in double. */
void or_matvec_f64(const float* A, int n, int ld, const double* x, const double* b, int nrhs, double* r) {
    for (int k = 0; k < nrhs; ++k) {
        const double* xv = x + (size_t)k * n;
        double* rv = r + (size_t)k * n;
        for (int i = 0; i < n; ++i) rv[i] = b[(size_t)k * n + i];
        for (int j = 0; j < n; ++j) {
            const double xj = xv[j];
            const float* col = A + (size_t)j * ld;
            for (int i = 0; i < n; ++i) rv[i] -= (double)col[i] * xj;
        }
    }
}

/* ------------------------------------------------------------------ matrix transpose */

void or_transpose_i32(const int32_t* in, int32_t* out, size_t rows, size_t cols) {
    const size_t n = rows * cols;
    if (!n) return;
    if (n == 1) { out[0] = in[0]; return; }
    for (size_t i = 0; i + 1 < n; ++i) out[i] = in[(size_t)(((unsigned __int128)i * cols) % (n - 1))];
    out[n - 1] = in[n - 1];
}

void or_transpose_blocks(const int32_t* in, int32_t* out, size_t rows, size_t cols, size_t bs) {
    const size_t bcc = cols / bs;
    const size_t subtasks = (rows / bs) * bcc;
    for (size_t s = 0; s < subtasks; ++s) {
        const size_t bir = s / bcc, bic = s % bcc;
        for (size_t i = 0; i < bs; ++i)          /* output row bic * bs + i, input column */
            for (size_t j = 0; j < bs; ++j)      /* output column bir * bs + j, input row */
                out[(bic * bs + i) * rows + bir * bs + j] = in[(bir * bs + j) * cols + bic * bs + i];
    }
}

/* ------------------------------------------------------------------ prefix sum */

void or_prefix_sum_i32(const int32_t* in, int32_t* out, size_t n, size_t chunk) {
    if (!n) return;
    if (!chunk) chunk = n;
    const size_t chunks = (n + chunk - 1) / chunk;
    /* per-chunk inclusive scans */
    for (size_t c = 0; c < chunks; ++c) {
        const size_t a = c * chunk, b = a + chunk < n ? a + chunk : n;
        uint32_t acc = 0;
        for (size_t i = a; i < b; ++i) { acc += (uint32_t)in[i]; out[i] = (int32_t)acc; }
    }
    /* scanned chunk totals added to the chunks after the first */
    for (size_t c = 1; c < chunks; ++c) {
        /* the previous chunk already holds the global prefix: its last element is the carry */
        const uint32_t carry = (uint32_t)out[c * chunk - 1];
        const size_t a = c * chunk, b = a + chunk < n ? a + chunk : n;
        for (size_t i = a; i < b; ++i) out[i] = (int32_t)((uint32_t)out[i] + carry);
    }
}

/* ------------------------------------------------------------------ sparse x dense */

void or_spmm_coo(const int32_t* rows, const int32_t* cols, const float* vals, size_t nnz, const float* B, size_t n,
                 float* Cf, double* Cd, int dbl) {
    if (dbl) memset(Cd, 0, n * n * sizeof(double));
    else memset(Cf, 0, n * n * sizeof(float));
    for (size_t k = 0; k < nnz; ++k) {
        const size_t r = (size_t)rows[k], c = (size_t)cols[k];
        for (size_t j = 0; j < n; ++j) {
            if (dbl) {
                Cd[r * n + j] += (double)vals[k] * (double)B[c * n + j];
            } else {
                volatile float p = vals[k] * B[c * n + j];   /* no fused multiply-add */
                Cf[r * n + j] = Cf[r * n + j] + p;
            }
        }
    }
}

/* ------------------------------------------------------------------ 2-D DFT */

/* One forward DFT of n = 2^k points in double, in place over x (interleaved re, im) read and
 * written with stride s (complex elements): bit reversal, then radix-2 butterflies with
 * w = exp(-2 pi i j / len) from cos / sin at each length. */
static void fft1_f64(double* x, size_t n, size_t s) {
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double* a = x + 2 * i * s;
            double* b = x + 2 * j * s;
            const double t0 = a[0], t1 = a[1];
            a[0] = b[0]; a[1] = b[1];
            b[0] = t0; b[1] = t1;
        }
    }
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t h = len / 2;
        for (size_t j = 0; j < h; ++j) {
            const double wr = cos(two_pi * (double)j / (double)len), wi = -sin(two_pi * (double)j / (double)len);
            for (size_t b = 0; b < n; b += len) {
                double* p = x + 2 * (b + j) * s;
                double* q = x + 2 * (b + j + h) * s;
                const double tr = wr * q[0] - wi * q[1], ti = wr * q[1] + wi * q[0];
                q[0] = p[0] - tr; q[1] = p[1] - ti;
                p[0] += tr; p[1] += ti;
            }
        }
    }
}

void or_fft2_f64(const float* in, double* out, size_t rows, size_t cols) {
    for (size_t i = 0; i < 2 * rows * cols; ++i) out[i] = (double)in[i];
    for (size_t r = 0; r < rows; ++r) fft1_f64(out + 2 * r * cols, cols, 1);      /* row transforms */
    /* column transforms, each on a contiguous copy of the column */
    double* col = (double*)malloc(2 * rows * sizeof(double));
    for (size_t c = 0; c < cols; ++c) {
        for (size_t r = 0; r < rows; ++r) {
            col[2 * r] = out[2 * (r * cols + c)];
            col[2 * r + 1] = out[2 * (r * cols + c) + 1];
        }
        fft1_f64(col, rows, 1);
        for (size_t r = 0; r < rows; ++r) {
            out[2 * (r * cols + c)] = col[2 * r];
            out[2 * (r * cols + c) + 1] = col[2 * r + 1];
        }
    }
    free(col);
}

/* ------------------------------------------------------------------ pageRank over an inverted index */

/* records: pages x (max_out + 1) uint32 (count, 1-based targets, padding). Every subtask s of P
 * pages fills its own window group by group from the top (invertIndex_cpu), then the groups are
 * concatenated order by order, subtasks ascending within an order (the redistribution).
 * entries: room for every link (3 uint32 each); counts[g]: entries of group g; returns the total. */
uint64_t or_prii_invert(const uint32_t* records, uint32_t pages, uint32_t max_out, uint32_t P, uint32_t* entries,
                        uint32_t* counts) {
    const uint32_t D = (pages + P - 1) / P;
    const uint64_t rec = (uint64_t)max_out + 1;
    uint64_t links = 0;
    for (uint64_t p = 0; p < pages; ++p) links += records[p * rec];
    uint32_t* win = (uint32_t*)malloc((links ? links : 1) * 3 * sizeof(uint32_t));
    uint64_t* per = (uint64_t*)calloc((size_t)D * D, sizeof(uint64_t));     /* per[s * D + g] */
    uint64_t* wbase = (uint64_t*)calloc(D + 1, sizeof(uint64_t));           /* window start of s */
    for (uint32_t s = 0; s < D; ++s) {
        const uint64_t p0 = (uint64_t)s * P, p1 = p0 + P < pages ? p0 + P : pages;
        uint64_t n = 0;
        for (uint64_t p = p0; p < p1; ++p)
            for (uint32_t k = 0; k < records[p * rec]; ++k) {
                ++per[(uint64_t)s * D + (records[p * rec + 1 + k] - 1) / P];
                ++n;
            }
        wbase[s + 1] = wbase[s] + n;
        /* top of each group, then fill downwards */
        uint64_t* loc = (uint64_t*)malloc((size_t)D * sizeof(uint64_t));
        uint64_t end = 0;
        for (uint32_t g = 0; g < D; ++g) { end += per[(uint64_t)s * D + g]; loc[g] = end; }
        uint32_t* w = win + 3 * wbase[s];
        for (uint64_t p = p0; p < p1; ++p)
            for (uint32_t k = 0; k < records[p * rec]; ++k) {
                const uint32_t t = records[p * rec + 1 + k];
                const uint64_t i = --loc[(t - 1) / P];
                w[3 * i] = (uint32_t)p;
                w[3 * i + 1] = t - 1;
                w[3 * i + 2] = records[p * rec];
            }
        free(loc);
    }
    uint64_t pos = 0;
    for (uint32_t g = 0; g < D; ++g) {
        uint64_t c = 0;
        for (uint32_t s = 0; s < D; ++s) {
            uint64_t off = 0;
            for (uint32_t h = 0; h < g; ++h) off += per[(uint64_t)s * D + h];
            const uint64_t n = per[(uint64_t)s * D + g];
            memcpy(entries + 3 * pos, win + 3 * (wbase[s] + off), n * 3 * sizeof(uint32_t));
            pos += n;
            c += n;
        }
        counts[g] = (uint32_t)c;
    }
    free(win);
    free(per);
    free(wbase);
    return pos;
}

/* `iterations` pageRank tasks over the inverted index (pageRank_cpu per subtask, entries in index
 * order): out[dest] += 0.85f * (x[src] / outlinks), x = 1.0 on the first iteration. */
void or_prii_pagerank(const uint32_t* entries, const uint32_t* counts, uint32_t pages, uint32_t P, int iterations,
                      float* out) {
    const uint32_t D = (pages + P - 1) / P;
    float* x = (float*)malloc((size_t)pages * sizeof(float));
    for (int it = 0; it < iterations; ++it) {
        memset(out, 0, (size_t)pages * sizeof(float));
        const uint32_t* e = entries;
        for (uint32_t g = 0; g < D; ++g)
            for (uint32_t i = 0; i < counts[g]; ++i, e += 3) {
                const float r = it == 0 ? 1.0f : x[e[0]];
                out[e[1]] += 0.85f * (r / (float)e[2]);
            }
        memcpy(x, out, (size_t)pages * sizeof(float));
    }
    free(x);
}

/* serialPageRank: pages in order push (0.85f * rank) / outlinks to every target. */
void or_prii_serial(const uint32_t* records, uint32_t pages, uint32_t max_out, int iterations, float* out) {
    const uint64_t rec = (uint64_t)max_out + 1;
    float* x = (float*)malloc((size_t)pages * sizeof(float));
    for (int it = 0; it < iterations; ++it) {
        memset(out, 0, (size_t)pages * sizeof(float));
        for (uint64_t p = 0; p < pages; ++p) {
            const uint32_t c = records[p * rec];
            const float incr = (float)(0.85f * (it == 0 ? 1.0f : x[p]) / (float)c);
            for (uint32_t k = 0; k < c; ++k) out[records[p * rec + 1 + k] - 1] += incr;
        }
        memcpy(x, out, (size_t)pages * sizeof(float));
    }
    free(x);
}
