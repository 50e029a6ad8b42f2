// fft kernels (replace cufftPlan1d / cufftPlanMany + cufftExecC2C of fft_cudaLaunchFunc, reference
// fftCUDA.cu:112-200): forward, unnormalised complex-to-complex DFTs of many sequences of
// n = 2^k single-precision complex points,
//   X[k] = sum_j x[j] exp(-2 pi i j k / n),
// the FFTW_FORWARD / CUFFT_FORWARD convention. Element j of sequence s sits at complex index
// s * seq_stride + j * elem_stride, so one entry point serves the row FFTs (elem_stride 1) and the
// column FFTs (seq_stride 1, elem_stride = row length) of the 2-D transform.
//
// gfx950 design. A workgroup transforms S sequences completely on chip: Stockham autosort passes
// (Govindaraju et al., "High performance discrete Fourier transforms on graphics processors",
// SC 2008) of radix 16 (n = 4096: three passes), preceded by one radix-2/4/8 pass when log2 n is
// not a multiple of 4. Each thread keeps its butterflies' points in registers and does the
// radix-16 DFT there as 4 x 4 (radix 8 as 2 x 4) with the inner twiddles as constants; the
// exchange between passes goes through LDS (n + n/16 complex per sequence: one pad slot per 16
// keeps the strided writes of the first passes off shared banks). Only the first pass reads HBM
// and only the last pass writes it, so a transform moves 16 bytes per point: the kernel is
// HBM-bound at 5 n log2 n flops per 16 n bytes.
//   - Row sequences (elem_stride 1): one workgroup per row for n >= 4096 (256 threads, one
//     radix-16 butterfly each per pass), 256 / (n/16) rows per workgroup below that.
//   - Column sequences (seq_stride 1): lanes run across adjacent columns first, so a wave reads
//     whole 32-128 byte row segments; S columns per workgroup (4 at n = 4096, 139 KB of LDS), and
//     the workgroups that share a line are placed on one XCD.
// Pass twiddles exp(-2 pi i m / 8192) come from one 64 KB table per device, computed in double on
// the host and rounded once to float (n = 8192 / 2^q reads every 2^q-th entry); the table is a
// device global filled at the first call on each device.
#include "uh_internal.h"

#include <cmath>
#include <mutex>

namespace uh {

static constexpr int FFT_MAX_LOG2 = 13;
static constexpr int FFT_TW = 1 << FFT_MAX_LOG2;          // twiddle table entries
static constexpr size_t FFT_LDS_CAP = 140 * 1024;          // LDS budget of one workgroup

__device__ float2 g_fft_tw[FFT_TW];

// HBM accesses go through global-space pointers to a plain two-float vector (HIP's float2 class
// cannot bind to an address-space-qualified object)
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// exp(-2 pi i t / 16) for t in [0, 16), from cos(pi/8), sin(pi/8) and sqrt(1/2).
#define UH_C8 0.92387953251128674f
#define UH_S8 0.38268343236508978f
#define UH_R2 0.70710678118654752f
__device__ __forceinline__ float2 w16(int t) {
    // cos and -sin of 2 pi t / 16
    constexpr float c[16] = {1.f, UH_C8, UH_R2, UH_S8, 0.f, -UH_S8, -UH_R2, -UH_C8,
                             -1.f, -UH_C8, -UH_R2, -UH_S8, 0.f, UH_S8, UH_R2, UH_C8};
    constexpr float s[16] = {0.f, -UH_S8, -UH_R2, -UH_C8, -1.f, -UH_C8, -UH_R2, -UH_S8,
                             0.f, UH_S8, UH_R2, UH_C8, 1.f, UH_C8, UH_R2, UH_S8};
    return make_float2(c[t], s[t]);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x * exp(-2 pi i e / R), e and R compile-time after unrolling
template <int R>
__device__ __forceinline__ float2 mulw(float2 x, int e) {
    const int t = (e % R) * (16 / R);
    if (t == 0) return x;
    if (t == 4) return make_float2(x.y, -x.x);       // -i
    if (t == 8) return make_float2(-x.x, -x.y);      // -1
    if (t == 12) return make_float2(-x.y, x.x);      // +i
    return cmul(x, w16(t));
}

template <int R> struct Dft;

template <> struct Dft<1> {
    __device__ __forceinline__ static void run(float2*) {}
};
template <> struct Dft<2> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    }
};
template <> struct Dft<4> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        // -i * d13 = (d13.y, -d13.x)
        v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
        v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
        v[1] = make_float2(d02.x + d13.y, d02.y - d13.x);
        v[3] = make_float2(d02.x - d13.y, d02.y + d13.x);
    }
};
// R = R1 * R2 with n = a + R1 m, k = b + R2 c:
//   X[b + R2 c] = sum_a W_R1^(a c) W_R^(a b) sum_m x[a + R1 m] W_R2^(m b)
template <int R1, int R2>
__device__ __forceinline__ void dft_split(float2* v) {
    constexpr int R = R1 * R2;
    float2 u[R1][R2];
#pragma unroll
    for (int a = 0; a < R1; ++a) {
        float2 w[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) w[m] = v[a + R1 * m];
        Dft<R2>::run(w);
#pragma unroll
        for (int b = 0; b < R2; ++b) u[a][b] = mulw<R>(w[b], a * b);
    }
#pragma unroll
    for (int b = 0; b < R2; ++b) {
        float2 w[R1];
#pragma unroll
        for (int a = 0; a < R1; ++a) w[a] = u[a][b];
        Dft<R1>::run(w);
#pragma unroll
        for (int c = 0; c < R1; ++c) v[b + R2 * c] = w[c];
    }
}
template <> struct Dft<8> {
    __device__ __forceinline__ static void run(float2* v) { dft_split<2, 4>(v); }
};
template <> struct Dft<16> {
    __device__ __forceinline__ static void run(float2* v) { dft_split<4, 4>(v); }
};

// Plan of a transform of 2^L points: one leading pass of radix 2^(L mod 4) (if nonzero), then
// radix-16 passes.
template <int L>
struct FftPlan {
    static constexpr int N = 1 << L;
    static constexpr int R0 = (L % 4) ? (1 << (L % 4)) : 16;
    static constexpr int PASSES = (L % 4 ? 1 : 0) + L / 4;
    static constexpr int T = N >= 4096 ? 256 : (N / 16 > 0 ? N / 16 : 1);   // threads per sequence
    static constexpr int PER = N / T;                                          // points per thread
    static constexpr int LDS_SEQ = N + N / 16;                                 // padded points
};

template <int L, int S, bool SEQ_FAST>
struct FftCfg {
    using P = FftPlan<L>;
    static constexpr int THREADS = P::T * S;
    static constexpr size_t LDS_BYTES = P::PASSES > 1 ? (size_t)P::LDS_SEQ * S * 8 : 0;
    __device__ __forceinline__ static int pad(int i) { return i + (i >> 4); }
    __device__ __forceinline__ static int lidx(int s, int i) {
        return SEQ_FAST ? pad(i) * S + s : s * P::LDS_SEQ + pad(i);
    }
};

// One Stockham pass of radix R over a sequence, butterflies b = t, t + T, ... (NS = product of the
// earlier radices). Reads: point b + r N/R. Writes: (b / NS) NS R + b % NS + r NS.
template <int L, int S, bool SEQ_FAST, int R, int NS, bool FIRST, bool LAST>
__device__ __forceinline__ void fft_pass(float2* lds, int s, int t, bool live, const UH_GLOBAL f32x2_t* in,
                                         uint64_t in_es, UH_GLOBAL f32x2_t* out, uint64_t out_es) {
    using P = FftPlan<L>;
    using Cf = FftCfg<L, S, SEQ_FAST>;
    constexpr int N = P::N;
    constexpr int NB = N / R / P::T;          // butterflies per thread
    static_assert(NB >= 1 && NB * R == P::PER, "plan");
    float2 v[NB][R];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = t + q * P::T;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = b + r * (N / R);
            if (FIRST) {
                const f32x2_t x = live ? in[(uint64_t)i * in_es] : f32x2_t{0.f, 0.f};
                v[q][r] = make_float2(x.x, x.y);
            } else {
                v[q][r] = lds[Cf::lidx(s, i)];
            }
        }
    }
    if (!FIRST && !LAST) __syncthreads();          // every read of this pass precedes its writes
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = t + q * P::T;
        if (NS > 1) {
            const int k = b % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], g_fft_tw[k * r * (FFT_TW / (NS * R))]);
        }
        Dft<R>::run(v[q]);
        const int d = (b / NS) * NS * R + b % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = d + r * NS;
            if (LAST) {
                if (live) out[(uint64_t)i * out_es] = f32x2_t{v[q][r].x, v[q][r].y};
            } else {
                lds[Cf::lidx(s, i)] = v[q][r];
            }
        }
    }
    if (!LAST) __syncthreads();
}

template <int L, int S, bool SEQ_FAST, int PASS, int NS>
__device__ __forceinline__ void fft_passes(float2* lds, int s, int t, bool live, const UH_GLOBAL f32x2_t* in,
                                           uint64_t in_es, UH_GLOBAL f32x2_t* out, uint64_t out_es) {
    using P = FftPlan<L>;
    constexpr int R = PASS == 0 ? P::R0 : 16;
    constexpr bool FIRST = PASS == 0, LAST = PASS == P::PASSES - 1;
    fft_pass<L, S, SEQ_FAST, R, NS, FIRST, LAST>(lds, s, t, live, in, in_es, out, out_es);
    if constexpr (!LAST) fft_passes<L, S, SEQ_FAST, PASS + 1, NS * R>(lds, s, t, live, in, in_es, out, out_es);
}

// Column groups whose S columns share 128-byte lines (K groups per line, S * 8 bytes each) run
// on one XCD: workgroup b runs on XCD b mod 8, so in every run of 8K workgroups, group
// 8K q + K r + u goes to workgroup 8K q + 8u + r. The K reads of a line then meet in that XCD's L2
// instead of each coming from HBM. A grid tail shorter than 8K keeps the identity order.
template <int S, bool SEQ_FAST>
__device__ __forceinline__ uint32_t fft_group(uint32_t b, uint32_t grid) {
    constexpr uint32_t K = (SEQ_FAST && S * 8 < 128) ? 128 / (S * 8) : 1;
    if constexpr (K == 1) {
        return b;
    } else {
        const uint32_t q = b / (8 * K);
        if ((q + 1) * 8 * K > grid) return b;
        const uint32_t w = b % (8 * K);
        return q * 8 * K + (w % 8) * K + w / 8;
    }
}

template <int L, int S, bool SEQ_FAST>
__global__ void __launch_bounds__((FftCfg<L, S, SEQ_FAST>::THREADS))
k_fft(const float2* in, uint64_t in_ss, uint64_t in_es, float2* out, uint64_t out_ss, uint64_t out_es,
      uint32_t count) {
    using P = FftPlan<L>;
    extern __shared__ float2 fft_lds[];
    const int tid = threadIdx.x;
    const int s = SEQ_FAST ? tid % S : tid / P::T;
    const int t = SEQ_FAST ? tid / S : tid % P::T;
    const uint64_t seq = (uint64_t)fft_group<S, SEQ_FAST>(blockIdx.x, gridDim.x) * S + s;
    const bool live = seq < count;
    const UH_GLOBAL f32x2_t* gi = (const UH_GLOBAL f32x2_t*)as_global(in) + (live ? seq * in_ss : 0);
    UH_GLOBAL f32x2_t* go = (UH_GLOBAL f32x2_t*)as_global(out) + (live ? seq * out_ss : 0);
    fft_passes<L, S, SEQ_FAST, 0, 1>(fft_lds, s, t, live, gi, in_es, go, out_es);
}

// ---- launch table: per log2 n, a row kernel (sequences slow) and a column kernel (fast)

typedef void (*FftKernel)(const float2*, uint64_t, uint64_t, float2*, uint64_t, uint64_t, uint32_t);
struct FftEntry {
    FftKernel k;
    int seqs;          // sequences per workgroup
    int threads;
    size_t lds;
};

template <int L>
constexpr int row_seqs() { return FftPlan<L>::T >= 256 ? 1 : 256 / FftPlan<L>::T; }
template <int L>
constexpr int col_seqs() {
    // lanes across >= 16 adjacent columns (128-byte segments) where LDS and 1024 threads allow
    constexpr int by_threads = 1024 / FftPlan<L>::T;
    constexpr int by_lds = (int)(FFT_LDS_CAP / ((size_t)FftPlan<L>::LDS_SEQ * 8));
    constexpr int m = by_threads < by_lds ? by_threads : by_lds;
    return m < 16 ? m : 16;
}

template <int L, bool COL>
FftEntry fft_entry() {
    constexpr int S = COL ? col_seqs<L>() : row_seqs<L>();
    static_assert(S >= 1, "sequences per workgroup");
    using Cf = FftCfg<L, S, COL>;
    static_assert(Cf::THREADS <= 1024 && Cf::LDS_BYTES <= FFT_LDS_CAP, "workgroup limits");
    return FftEntry{k_fft<L, S, COL>, S, Cf::THREADS, Cf::LDS_BYTES};
}

template <int... Ls>
struct FftTable {
    static FftEntry get(int l, bool col) {
        FftEntry e{nullptr, 0, 0, 0};
        ((l == Ls ? (e = col ? fft_entry<Ls, true>() : fft_entry<Ls, false>(), 0) : 0), ...);
        return e;
    }
};
using FftAll = FftTable<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13>;

// Twiddle table of each device, filled once (a synchronous copy, ordered before every later
// launch on any stream of the device). Kernels whose LDS exceeds 64 KB get their limit raised
// at the same time.
static std::mutex g_fft_mtx;
static bool g_fft_ready[64];

static int fft_prepare() {
    int dev = 0;
    UH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(UH_ERR_ARG, "uh_fft_c2c_f32: device index %d", dev);
    std::lock_guard<std::mutex> g(g_fft_mtx);
    if (g_fft_ready[dev]) return UH_OK;
    std::vector<float2> tw(FFT_TW);
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < FFT_TW; ++m) {
        const double a = two_pi * (double)m / (double)FFT_TW;
        tw[m] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    UH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_fft_tw), tw.data(), sizeof(float2) * FFT_TW, 0, hipMemcpyHostToDevice));
    for (int l = 1; l <= FFT_MAX_LOG2; ++l)
        for (int c = 0; c < 2; ++c) {
            const FftEntry e = FftAll::get(l, c != 0);
            if (e.lds > 64 * 1024)
                UH_HIP(hipFuncSetAttribute((const void*)e.k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds));
        }
    g_fft_ready[dev] = true;
    return UH_OK;
}

}  // namespace uh

using namespace uh;

extern "C" int uh_fft_c2c_f32(const void* in, uint64_t in_seq_stride, uint64_t in_elem_stride, void* out,
                              uint64_t out_seq_stride, uint64_t out_elem_stride, uint32_t n, uint32_t count,
                              void* stream) {
    if (!count) return UH_OK;
    if (!in || !out) return fail(UH_ERR_ARG, "uh_fft_c2c_f32: null buffer");
    if (n < 2 || n > (uint32_t)FFT_TW || (n & (n - 1)))
        return fail(UH_ERR_ARG, "uh_fft_c2c_f32: n = %u is not a power of two in [2, %d]", n, FFT_TW);
    if (((uintptr_t)in | (uintptr_t)out) & 7u) return fail(UH_ERR_ARG, "uh_fft_c2c_f32: buffers must be 8-byte aligned");
    if (!in_elem_stride || !out_elem_stride) return fail(UH_ERR_ARG, "uh_fft_c2c_f32: zero element stride");
    if (in == out && (in_seq_stride != out_seq_stride || in_elem_stride != out_elem_stride))
        return fail(UH_ERR_ARG, "uh_fft_c2c_f32: in-place transforms need equal strides");
    int rc = fft_prepare();
    if (rc) return rc;
    hipStream_t s = plain_stream_of(stream, &rc);
    if (rc) return rc;
    int l = 0;
    while ((1u << l) < n) ++l;
    // column layout when adjacent sequences are adjacent in memory
    const bool col = in_seq_stride == 1 && out_seq_stride == 1 && in_elem_stride > 1;
    const FftEntry e = FftAll::get(l, col);
    const uint64_t groups = ((uint64_t)count + e.seqs - 1) / e.seqs;
    UH_LAUNCH(e.k, dim3((unsigned)groups), dim3(e.threads), e.lds, s, (const float2*)in, in_seq_stride, in_elem_stride,
              (float2*)out, out_seq_stride, out_elem_stride, count);
    UH_LAUNCH_CHECK("k_fft");
    g_plain_launches++;
    return UH_OK;
}
