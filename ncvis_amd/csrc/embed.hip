// Layout initialisation, NCE SGD epochs and the explicit NCE loss / adjoint on gfx950.
//
// Every data set of a fit is one FitSeg descriptor in device memory; the init and epoch kernels
// walk the concatenation of all data sets in one launch (a thread finds its data set by binary
// search over the descriptors' offsets), so one fit and a batch of many share one code path and
// a batch fills the GPU with all its data sets' rows and chains at once. Edge and chain counts
// are read from the device (k_seg_dims), so no stage waits for the host.
//
// Initialisation (src/ncvis.cpp:184-273):
//   k_init_random_seg  Y[i] = uniform_real<float>(0,1) drawn from pcg64(seed + t), where t is
//                      the thread that owns element i under libgomp's static schedule of the N*d
//                      loop (ncvis.cpp:198-201); a GPU thread jumps the stream to its element.
//   k_spmm_seg         Y_new[i] = sum_{edges (i, j) in row order} alpha * Y_old[j], row-owned
//                      (race-free), uncontracted like the reference's loop (ncvis.cpp:206-218);
//                      a wave gathers its 64 rows' neighbour rows into LDS, each lane sums its own
//   k_colstat_seg / k_finalize_seg / k_normalize_seg
//                      per-dimension mean and population std (FP64 partial sums over a fixed
//                      block partition, deterministic) and standardisation (ncvis.cpp:220-260)
// Epochs (src/ncvis.cpp:275-352):
//   k_sgd_epoch_seg    chain t of a data set = the reference's OpenMP thread t: pcg64(seed + t),
//                      the static edge chunk [s0, e0) of thread t, a private Q copy. Mode 0 is
//                      the exact form (plain loads/stores, one sample at a time; one chain is
//                      the reference at one thread); mode 1 the same with float atomics; mode 2
//                      the many-chain form (source row in registers with the chain's own
//                      updates, far endpoints by 64-bit compare-and-swap per coordinate pair).
//   k_q_reduce_seg     Q = mean of the chains' Q copies (ncvis.cpp:342-349), FP64 sum.
// NCE loss (the objective the update of ncvis.cpp:304-338 descends):
//   k_nce_loss         per edge: positive + n explicit negatives; loss and dL/dQ as FP64 partial
//                      sums, dL/dY scattered with FP32 atomics.
#include <algorithm>
#include <cfloat>
#include <stdexcept>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

// ------------------------------------------------------------------------ descriptors
__global__ void k_seg_set(FitSeg *__restrict__ dst, FitSeg v) { *dst = v; }

void seg_set(FitSeg *d_seg, const FitSeg &v, hipStream_t st) {
    k_seg_set<<<1, 1, 0, st>>>(d_seg, v);
    NCVIS_LAUNCH_CHECK();
}

// data set holding global position x of the offsets selected by OFF (0: rows, 1: chain slots)
template <int OFF>
__device__ __forceinline__ int seg_find(const FitSeg *__restrict__ s, int B, int64_t x) {
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        const int64_t o = OFF == 0 ? s[m].row_off : s[m].chain_off;
        if (o <= x)
            lo = m;
        else
            hi = m;
    }
    return lo;
}

// E = row_ptr[N], T by the chain rule, and the epoch mode of every data set
__global__ void k_seg_dims(FitSeg *__restrict__ s, int B, ChainRule rule, int d) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    FitSeg &g = s[b];
    const int64_t E = g.row_ptr[g.N];
    int64_t T = chain_rule(E, rule);
    if (T > g.T_cap) T = g.T_cap;
    const int atomic = rule.atomic_opt >= 0 ? rule.atomic_opt : (T > 1 ? 1 : 0);
    g.E = E;
    g.T = T;
    int mode = atomic ? ((d >= 1 && d <= 4 && (d % 2 != 0 || g.pairs_ok)) ? 2 : 1) : 0;
    if (atomic && rule.replicas && g.Yrep && d >= 1 && d <= 4) mode = 3;
    g.mode = mode;
}

void seg_dims(FitSeg *d_segs, int B, const ChainRule &rule, int d, hipStream_t st) {
    k_seg_dims<<<cdiv(B, 64), 64, 0, st>>>(d_segs, B, rule, d);
    NCVIS_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------ init
// Elements [i0, i0 + 64) of the concatenated N*d element space; a run may span data sets.
__global__ void k_init_random_seg(const FitSeg *__restrict__ s, int B, int64_t total, int d, uint64_t seed) {
    constexpr int RUN = 64;
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * RUN;
    if (i0 >= total) return;
    const int64_t i1 = i0 + RUN < total ? i0 + RUN : total;
    int64_t i = i0;
    while (i < i1) {
        const int b = seg_find<0>(s, B, i / d);
        const FitSeg &g = s[b];
        const int64_t base = g.row_off * d, Nd = g.N * d, T = g.T;
        const int64_t l0 = i - base;
        const int64_t l1 = (i1 - base < Nd) ? i1 - base : Nd;
        // owner of element l0 under the static schedule of Nd over T threads
        const int64_t q = Nd / T, r = Nd % T;
        int64_t t = (l0 < r * (q + 1)) ? l0 / (q + 1) : r + (l0 - r * (q + 1)) / q;
        int64_t s0, e0;
        static_chunk(Nd, T, t, s0, e0);
        U128 st = pcg_advance(pcg_seed(seed + (uint64_t)t), (uint64_t)(l0 - s0));
        for (int64_t l = l0; l < l1; ++l) {
            while (l >= e0) {
                ++t;
                static_chunk(Nd, T, t, s0, e0);
                st = pcg_seed(seed + (uint64_t)t);
            }
            g.Y[l] = uniform_unit_float(st);
        }
        i = base + l1;
    }
}

// One row of the SpMM by one thread, in edge order: each edge's index and row read once, four
// edges fetched ahead of their in-order accumulation.
template <int DIM>
__device__ __forceinline__ void spmm_row(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                         float alpha, const float *__restrict__ Yo, float *__restrict__ Yn,
                                         int64_t i) {
    float acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = 0.f;
    const int64_t b = row_ptr[i], e = row_ptr[i + 1];
    int64_t p = b;
    for (; p + 4 <= e; p += 4) {
        int32_t j[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = col[p + u];
        float y[4][DIM];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < DIM; ++c) y[u][c] = Yo[(int64_t)j[u] * DIM + c];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < DIM; ++c) acc[c] = __fadd_rn(acc[c], __fmul_rn(alpha, y[u][c]));
    }
    for (; p < e; ++p) {
        const int64_t j = col[p];
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] = __fadd_rn(acc[c], __fmul_rn(alpha, Yo[j * DIM + c]));
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) Yn[i * DIM + c] = acc[c];
}

// Parity 0 reads Y and writes Y2, 1 the reverse. Every dimension's sum keeps the sequential edge
// order (the reference's), so all forms below give the same bits.
// DIM 0: any d, one thread per row, one dimension at a time.
// DIM 1..4: a wave takes 64 consecutive rows of one data set, whose edges form one contiguous
// range: it reads their column indices coalesced and gathers the neighbour rows 64 at a time into
// LDS, 256 edges per round, and each lane then adds up the part of its own row in that round. A
// wave whose rows span two data sets runs one thread per row (spmm_row).
constexpr int kSpmmChunk = 256;  // edges per wave and round

template <int DIM>
__global__ __launch_bounds__(256) void k_spmm_seg(const FitSeg *__restrict__ s, int B, int64_t R, int d_rt,
                                                  int parity) {
    if constexpr (DIM == 0) {
        const int64_t gr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (gr >= R) return;
        const FitSeg &g = s[seg_find<0>(s, B, gr)];
        const int64_t i = gr - g.row_off;
        const float *__restrict__ Yo = parity ? g.Y2 : g.Y;
        float *__restrict__ Yn = parity ? g.Y : g.Y2;
        const int64_t b = g.row_ptr[i], e = g.row_ptr[i + 1];
        for (int c = 0; c < d_rt; ++c) {
            float acc = 0.f;
            for (int64_t p = b; p < e; ++p)
                acc = __fadd_rn(acc, __fmul_rn(g.alpha_init, Yo[(int64_t)g.col[p] * d_rt + c]));
            Yn[i * d_rt + c] = acc;
        }
    } else {
        __shared__ float ly[4][kSpmmChunk * DIM];
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int64_t w0 = (int64_t)blockIdx.x * 256 + wave * 64;  // the wave's first row
        if (w0 >= R) return;                                       // whole wave; no block barrier below
        const int64_t wl = (w0 + 63 < R) ? w0 + 63 : R - 1;        // its last row
        const int sb = seg_find<0>(s, B, w0);
        const int64_t gr = w0 + lane;
        if (sb != seg_find<0>(s, B, wl)) {  // uniform
            if (gr < R) {
                const FitSeg &g = s[seg_find<0>(s, B, gr)];
                spmm_row<DIM>(g.row_ptr, g.col, g.alpha_init, parity ? g.Y2 : g.Y, parity ? g.Y : g.Y2,
                              gr - g.row_off);
            }
            return;
        }
        const FitSeg &g = s[sb];
        const float *__restrict__ Yo = parity ? g.Y2 : g.Y;
        float *__restrict__ Yn = parity ? g.Y : g.Y2;
        const int32_t *__restrict__ col = g.col;
        const float alpha = g.alpha_init;
        const int64_t i0 = w0 - g.row_off, nrow = wl - w0 + 1;
        const int64_t eb = g.row_ptr[i0], ee = g.row_ptr[i0 + nrow];  // the wave's edges
        const bool live = lane < nrow;
        const int64_t b = live ? g.row_ptr[i0 + lane] : 0, e = live ? g.row_ptr[i0 + lane + 1] : 0;
        float *my = ly[wave];
        float acc[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] = 0.f;
        constexpr int V = kSpmmChunk / 64;
        for (int64_t c0 = eb; c0 < ee; c0 += kSpmmChunk) {
            int32_t j[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const int64_t p = c0 + v * 64 + lane;
                j[v] = (p < ee) ? col[p] : -1;
            }
            float y[V][DIM];
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int c = 0; c < DIM; ++c) y[v][c] = (j[v] >= 0) ? Yo[(int64_t)j[v] * DIM + c] : 0.f;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int c = 0; c < DIM; ++c) my[(v * 64 + lane) * DIM + c] = y[v][c];
            // the round's rows are in LDS for every lane of the wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int64_t pb = (b > c0) ? b : c0;
            const int64_t pe = (e < c0 + kSpmmChunk) ? e : c0 + kSpmmChunk;
            for (int64_t p = pb; p < pe; ++p) {
#pragma unroll
                for (int c = 0; c < DIM; ++c)
                    acc[c] = __fadd_rn(acc[c], __fmul_rn(alpha, my[(int)(p - c0) * DIM + c]));
            }
            // every lane's reads of this round precede the next round's writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        if (live) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) Yn[(i0 + lane) * DIM + c] = acc[c];
        }
    }
}

// Per-block FP64 column sums of the freshly written buffer (MODE 0) or of (Y - mean)^2 (MODE 1).
// Grid (max nb, B): data set blockIdx.y uses its first nb blocks, each striding over its rows.
template <int MODE>
__global__ __launch_bounds__(256) void k_colstat_seg(const FitSeg *__restrict__ s, int d, int parity) {
    __shared__ double red[256];
    const FitSeg &g = s[blockIdx.y];
    if ((int)blockIdx.x >= g.nb) return;
    const float *__restrict__ Y = parity ? g.Y : g.Y2;
    const int64_t N = g.N;
    for (int c = 0; c < d; ++c) {
        double acc = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)g.nb * blockDim.x) {
            float v = Y[i * d + c];
            if (MODE == 1) {
                float dl = __fsub_rn(v, g.stats[c]);
                v = __fmul_rn(dl, dl);
            }
            acc += (double)v;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) g.part[(int64_t)blockIdx.x * d + c] = red[0];
        __syncthreads();
    }
}

// One 256-thread block per data set. MODE 0: mean[c] = sum / N; MODE 1: sigma[c] = sqrtf(sum / N).
// The block partials are summed in a fixed order (thread t takes partials t, t + 256, ..., then a
// fixed tree), so the statistics do not depend on scheduling.
template <int MODE>
__global__ __launch_bounds__(256) void k_finalize_seg(const FitSeg *__restrict__ s, int d) {
    __shared__ double red[256];
    const FitSeg &g = s[blockIdx.x];
    for (int c = 0; c < d; ++c) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < g.nb; b += 256) acc += g.part[(int64_t)b * d + c];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const float v = (float)(red[0] / (double)g.N);
            g.stats[MODE * d + c] = (MODE == 0) ? v : sqrtf(v);
        }
        __syncthreads();
    }
}

__global__ void k_normalize_seg(const FitSeg *__restrict__ s, int B, int64_t total, int d, int parity) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const FitSeg &g = s[seg_find<0>(s, B, e / d)];
    const int64_t l = e - g.row_off * d;
    const int c = (int)(l % d);
    float *Y = parity ? g.Y : g.Y2;
    Y[l] = __fsub_rn(Y[l], g.stats[c]) / g.stats[d + c];
}

__global__ void k_copy_back_seg(const FitSeg *__restrict__ s, int B, int64_t total, int d) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const FitSeg &g = s[seg_find<0>(s, B, e / d)];
    const int64_t l = e - g.row_off * d;
    g.Y[l] = g.Y2[l];
}

void seg_init_embedding(const FitSeg *d_segs, int B, const SegTotals &tot, int d, int n_init_epochs, uint64_t seed,
                        hipStream_t st) {
    const int64_t total = tot.rows * d;
    if (total <= 0) return;
    k_init_random_seg<<<cdiv(cdiv(total, 64), 256), 256, 0, st>>>(d_segs, B, total, d, seed);
    NCVIS_LAUNCH_CHECK();
    const dim3 sgrid((unsigned)std::max(1, tot.max_nb), (unsigned)B);
    for (int ep = 0; ep < n_init_epochs; ++ep) {
        const int par = ep % 2;  // epoch ep reads buffer par (0: Y) and writes the other
        switch (d) {
            case 1: k_spmm_seg<1><<<cdiv(tot.rows, 256), 256, 0, st>>>(d_segs, B, tot.rows, d, par); break;
            case 2: k_spmm_seg<2><<<cdiv(tot.rows, 256), 256, 0, st>>>(d_segs, B, tot.rows, d, par); break;
            case 3: k_spmm_seg<3><<<cdiv(tot.rows, 256), 256, 0, st>>>(d_segs, B, tot.rows, d, par); break;
            case 4: k_spmm_seg<4><<<cdiv(tot.rows, 256), 256, 0, st>>>(d_segs, B, tot.rows, d, par); break;
            default: k_spmm_seg<0><<<cdiv(tot.rows, 256), 256, 0, st>>>(d_segs, B, tot.rows, d, par); break;
        }
        NCVIS_LAUNCH_CHECK();
        k_colstat_seg<0><<<sgrid, 256, 0, st>>>(d_segs, d, par);
        k_finalize_seg<0><<<B, 256, 0, st>>>(d_segs, d);
        k_colstat_seg<1><<<sgrid, 256, 0, st>>>(d_segs, d, par);
        k_finalize_seg<1><<<B, 256, 0, st>>>(d_segs, d);
        k_normalize_seg<<<cdiv(total, 256), 256, 0, st>>>(d_segs, B, total, d, par);
        NCVIS_LAUNCH_CHECK();
    }
    if (n_init_epochs % 2) {
        k_copy_back_seg<<<cdiv(total, 256), 256, 0, st>>>(d_segs, B, total, d);
        NCVIS_LAUNCH_CHECK();
    }
}

static int stat_blocks(int64_t N) { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(N, 256))); }

void init_embedding_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, float alpha,
                        int n_init_epochs, uint64_t seed, int64_t chains, float *dY, hipStream_t st) {
    if (N < 1 || d < 1 || d > 1024) throw HipError("[ncvis_amd] init: invalid sizes");
    const int nb = stat_blocks(N);
    DevBuf<FitSeg> seg(1);
    DevBuf<float> y2(n_init_epochs > 0 ? (size_t)N * d : 0);
    DevBuf<double> part((size_t)nb * d);
    DevBuf<float> stats((size_t)2 * d);
    FitSeg v{};
    v.Y = dY;
    v.Y2 = y2.p;
    v.row_ptr = d_row_ptr;
    v.col = d_col;
    v.N = N;
    v.part = part.p;
    v.stats = stats.p;
    v.alpha_init = alpha;
    v.nb = nb;
    v.T = chains < 1 ? 1 : chains;
    v.T_cap = v.T;
    seg_set(seg.p, v, st);
    SegTotals tot;
    tot.rows = N;
    tot.slots = v.T;
    tot.max_nb = nb;
    seg_init_embedding(seg.p, 1, tot, d, n_init_epochs, seed, st);
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------------ SGD
// powf / expf evaluated in double and rounded once: correctly rounded in all but vanishingly
// rare cases, which is what the host libm's float versions return almost always.
__device__ __forceinline__ float pow_rn(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float exp_rn(float x) { return (float)exp((double)x); }

struct SampleConst {
    float a, b, alpha_Q, step;
    float noise_f;     // (float)cur_noise
    uint32_t noise;    // cur_noise
    int fast_math;     // 1: single-precision powf / expf; 2: hardware log2 / exp2 / reciprocal
    int edge_flush;    // mode 3: write the source row's accumulated update after every edge
    uint32_t rot;      // mode 3 with permuted copies: launch count that rotates the negatives' eighths
    int reread;        // mode 3: 1 read the far row again right before its store, 0 never (store the
                       // prefetched row minus the update), -1 unless the copies are permuted
};

// The sample body of ncvis.cpp:304-339 on a register copy of the two rows. Returns the weight
// applied; updates yi / yo in place and the chain's Q copy.
template <int DIM>
__device__ __forceinline__ void sgd_math(float (&yi)[DIM], float (&yo)[DIM], bool positive, const SampleConst &c,
                                         float &Qc, float (&dx)[DIM]) {
    float d2 = 0.f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        float t = __fsub_rn(yi[k], yo[k]);
        d2 = __fadd_rn(d2, __fmul_rn(t, t));
    }
    if (c.fast_math == 2) {
        // hardware log2 / exp2 / reciprocal (about 1 ulp each): d2^b and d2^(b-1) share one log2,
        // e^Q = 2^(Q log2 e). d2 = 0 gives log2 = -inf, so d2^b = 0 and d2^(b-1) = 0 or +inf as
        // powf gives; b = 1 keeps d2^0 = 1
        const float lg = __builtin_amdgcn_logf(d2);
        const float Ph = __builtin_amdgcn_rcpf(__fadd_rn(1.f, __fmul_rn(c.a, __builtin_amdgcn_exp2f(c.b * lg))));
        float w = 1.f;
        if (c.noise != 0) {
            w = Ph * __builtin_amdgcn_rcpf(c.noise_f * __builtin_amdgcn_exp2f(Qc * 1.44269504088896341f));
            if (positive)
                w = __builtin_amdgcn_rcpf(1.f + w);
            else
                w = -__builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_rcpf(w));
            Qc = __fsub_rn(Qc, __fmul_rn(w, c.alpha_Q));
            const float pb1 = c.b == 1.f ? 1.f : __builtin_amdgcn_exp2f((c.b - 1.f) * lg);
            w = 2.f * w * Ph * c.a * c.b * pb1;
        }
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            float v = __fmul_rn(__fmul_rn(__fsub_rn(yo[k], yi[k]), w), c.step);
            if (v > 4.f)
                v = 4.f;
            else if (v < -4.f)
                v = -4.f;
            dx[k] = v;
        }
        return;
    }
    float Ph = 1.f / __fadd_rn(1.f, __fmul_rn(c.a, c.fast_math ? powf(d2, c.b) : pow_rn(d2, c.b)));
    float w = 1.f;
    if (c.noise != 0) {
        w = Ph / (c.noise_f * (c.fast_math ? expf(Qc) : exp_rn(Qc)));
        if (positive)
            w = 1.f / (1.f + w);
        else
            w = -1.f / (1.f + 1.f / w);
        Qc = __fsub_rn(Qc, __fmul_rn(w, c.alpha_Q));
        w = 2.f * w * Ph * c.a * c.b * (c.fast_math ? powf(d2, c.b - 1.f) : pow_rn(d2, c.b - 1.f));
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        float v = __fmul_rn(__fmul_rn(__fsub_rn(yo[k], yi[k]), w), c.step);
        if (v > 4.f)
            v = 4.f;
        else if (v < -4.f)
            v = -4.f;
        dx[k] = v;
    }
}

template <int DIM>
__device__ __forceinline__ void sgd_sample(float *__restrict__ Y, int64_t id, int64_t other, bool positive,
                                           const SampleConst &c, float &Qc, int atomic_mode) {
    float yi[DIM], yo[DIM], dx[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        yi[k] = Y[id * DIM + k];
        yo[k] = Y[other * DIM + k];
    }
    sgd_math<DIM>(yi, yo, positive, c, Qc, dx);
    if (atomic_mode) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            __hip_atomic_fetch_add(&Y[id * DIM + k], dx[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&Y[other * DIM + k], -dx[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else {
        // Y[id] += dx; Y[other] -= dx, in the reference's order (other may equal id)
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            float ni = __fadd_rn(yi[k], dx[k]);
            Y[id * DIM + k] = ni;
            float cur = (other == id) ? ni : yo[k];
            Y[other * DIM + k] = __fsub_rn(cur, dx[k]);
        }
    }
}

// Generic-d variant working directly on memory (d not known at compile time).
__device__ inline void sgd_sample_rt(float *__restrict__ Y, int d, int64_t id, int64_t other, bool positive,
                                     const SampleConst &c, float &Qc, int atomic_mode) {
    float d2 = 0.f;
    for (int k = 0; k < d; ++k) {
        float t = __fsub_rn(Y[id * d + k], Y[other * d + k]);
        d2 = __fadd_rn(d2, __fmul_rn(t, t));
    }
    float Ph = 1.f / __fadd_rn(1.f, __fmul_rn(c.a, pow_rn(d2, c.b)));
    float w = 1.f;
    if (c.noise != 0) {
        w = Ph / (c.noise_f * exp_rn(Qc));
        if (positive)
            w = 1.f / (1.f + w);
        else
            w = -1.f / (1.f + 1.f / w);
        Qc = __fsub_rn(Qc, __fmul_rn(w, c.alpha_Q));
        w = 2.f * w * Ph * c.a * c.b * pow_rn(d2, c.b - 1.f);
    }
    for (int k = 0; k < d; ++k) {
        float v = __fmul_rn(__fmul_rn(__fsub_rn(Y[other * d + k], Y[id * d + k]), w), c.step);
        if (v > 4.f)
            v = 4.f;
        else if (v < -4.f)
            v = -4.f;
        if (atomic_mode) {
            __hip_atomic_fetch_add(&Y[id * d + k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&Y[other * d + k], -v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            Y[id * d + k] = __fadd_rn(Y[id * d + k], v);
            Y[other * d + k] = __fsub_rn(Y[other * d + k], v);
        }
    }
}


// ------------------------------------------------------------------------ epoch bodies
// chain_row[t] = row holding edge s0 of chain t (row_ptr[r] <= s0 < row_ptr[r+1])
__global__ void k_chain_rows_seg(const FitSeg *__restrict__ s, int B, int64_t slots, uint64_t seed) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= slots) return;
    const FitSeg &g = s[seg_find<1>(s, B, gt)];
    const int64_t t = gt - g.chain_off;
    if (t >= g.T) return;
    int64_t s0, e0;
    static_chunk(g.E, g.T, t, s0, e0);
    int64_t lo = 0, hi = g.N;  // largest r with row_ptr[r] <= s0
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (g.row_ptr[mid] <= s0)
            lo = mid;
        else
            hi = mid;
    }
    while (lo < g.N - 1 && g.row_ptr[lo + 1] <= s0) ++lo;
    g.chain_row[t] = lo;
    g.rng[t] = pcg_seed(seed + (uint64_t)t);
}

// One chain, sample by sample on memory (modes 0 and 1): the reference's thread loop
// (ncvis.cpp:290-340). DIM 0: d only known at run time.
template <int DIM>
__device__ __forceinline__ void chain_exact(float *__restrict__ Y, int d_rt, const int64_t *__restrict__ row_ptr,
                                            const int32_t *__restrict__ col, int64_t s0, int64_t e0, int64_t &r,
                                            U128 &g, float &Qc, const SampleConst &c, uint64_t N, int atomic_mode) {
    int64_t rend = row_ptr[r + 1];
    for (int64_t e = s0; e < e0; ++e) {
        while (e >= rend) {
            ++r;
            rend = row_ptr[r + 1];
        }
        const int64_t id = r;
        for (uint32_t j = 0; j <= c.noise; ++j) {
            int64_t other;
            if (j == 0) {
                other = col[e];
            } else {
                do {
                    other = (int64_t)uniform_index(g, N);
                } while (other == id);
            }
            if constexpr (DIM > 0)
                sgd_sample<DIM>(Y, id, other, j == 0, c, Qc, atomic_mode);
            else
                sgd_sample_rt(Y, d_rt, id, other, j == 0, c, Qc, atomic_mode);
        }
    }
}

// Row access of the per-XCD copies (mode 3): relaxed atomic loads and stores at workgroup scope
// (plain loads/stores that skip the CU's L1, so a row is read from the XCD's L2, where every
// update of that copy lands), coordinate pairs as one 8-byte access.
template <int DIM>
__device__ __forceinline__ void rep_load(const float *p, float (&v)[DIM]) {
    if constexpr (DIM % 2 == 0) {
#pragma unroll
        for (int k = 0; k < DIM; k += 2) {
            const unsigned long long u = __hip_atomic_load((const unsigned long long *)(p + k), __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            const float2 f = __builtin_bit_cast(float2, u);
            v[k] = f.x;
            v[k + 1] = f.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < DIM; ++k) v[k] = __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <int DIM>
__device__ __forceinline__ void rep_store(float *p, const float (&v)[DIM]) {
    if constexpr (DIM % 2 == 0) {
#pragma unroll
        for (int k = 0; k < DIM; k += 2)
            __hip_atomic_store((unsigned long long *)(p + k),
                               __builtin_bit_cast(unsigned long long, make_float2(v[k], v[k + 1])), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
#pragma unroll
        for (int k = 0; k < DIM; ++k) __hip_atomic_store(p + k, v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Hardware id of the XCD (L2 domain) this wave runs on.
__device__ __forceinline__ int xcd_id() {
    return __builtin_amdgcn_s_getreg((20 /* HW_REG_XCC_ID */) | (0 << 6) | (15 << 11)) & (kXcds - 1);
}

// One chain of the many-chain (Hogwild) modes. The chain keeps the row of its current source
// point in registers and adds its own updates of that row there (it sees its own writes, as a
// thread of the reference does), flushing them to Y when it moves to the next row; updates of the
// other endpoint go to Y at once. The negatives of an edge are drawn before their rows are read
// (the draw sequence does not depend on Y), so PF rows are fetched together instead of one
// dependent load per sample.
//   REP = false (mode 2): Y is shared by all XCDs; updates are device-scope atomics (a 64-bit
//     compare-and-swap per coordinate pair, float atomics for odd d), performed at the memory
//     side, so concurrent chains lose no update.
//   REP = true (mode 3): Y is this XCD's own copy, touched by no other XCD, so plain
//     read-modify-writes complete in the XCD's L2; the far row is read again right before it is
//     written, so only a chain of the same XCD updating that row in between loses its update
//     (the reference's Hogwild race, narrowed to one L2 round trip). The copies are merged after
//     every sub-epoch (k_rep_sync).
//   pa != 0 (mode 3 with negative slices): the copies hold row r at position (pa r) mod N, an
//     affine bijection that scatters any run of rows (a cluster of rows sorted by class, say)
//     evenly over the positions; a launch draws its negatives on an XCD as positions of one
//     eighth of the copy, which stays in that XCD's L2, every eighth in turn (c.rot), so each
//     negative is a uniform row over the launches and independent of the input order.
template <int DIM, int PF, bool REP>
__device__ __forceinline__ void chain_hw(float *__restrict__ Y, const int64_t *__restrict__ row_ptr,
                                         const int32_t *__restrict__ col, int64_t s0, int64_t e0, int64_t &r,
                                         U128 &g, float &Qc, const SampleConst &c, uint64_t N, int plain_far,
                                         uint64_t pa = 0) {
    // position of a row in Y (the identity unless the copies are permuted)
    auto pos = [&](int64_t row) -> int64_t {
        if constexpr (REP) {
            if (pa) return (int64_t)(((uint64_t)row * pa) % N);
        }
        return row;
    };
    int64_t rend = row_ptr[r + 1];
    int64_t rp = pos(r);  // position of the source row
    float base[DIM], acc[DIM];
    auto flush = [&](int64_t row) {
        if constexpr (REP) {
            bool any = false;
#pragma unroll
            for (int k = 0; k < DIM; ++k) any |= acc[k] != 0.f;
            if (any) {
                float cur[DIM];
                rep_load<DIM>(Y + row * DIM, cur);
#pragma unroll
                for (int k = 0; k < DIM; ++k) cur[k] = __fadd_rn(cur[k], acc[k]);
                rep_store<DIM>(Y + row * DIM, cur);
            }
        } else {
#pragma unroll
            for (int k = 0; k < DIM; ++k)
                if (acc[k] != 0.f)
                    __hip_atomic_fetch_add(&Y[row * DIM + k], acc[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    auto fetch = [&](int64_t row, float (&v)[DIM]) {
        if constexpr (REP) {
            rep_load<DIM>(Y + row * DIM, v);
        } else {
#pragma unroll
            for (int k = 0; k < DIM; ++k) v[k] = Y[row * DIM + k];
        }
    };
    fetch(rp, base);
#pragma unroll
    for (int k = 0; k < DIM; ++k) acc[k] = 0.f;
    const uint32_t total = c.noise + 1;
    // negatives are positions in [nlo, nlo + nrange): all of Y, or with permuted copies the
    // eighth given by this XCD and launch
    uint64_t nlo = 0, nrange = N;
    // the far row is read again right before its store (one L2 round trip of race window), except
    // on permuted copies: there the store writes the row as prefetched minus the update, the
    // reference's own window (Y[other] read at the sample's start, ncvis.cpp:306-336), which
    // saves a third of the L2 requests of a sample
    const bool reread = c.reread > 0 || (c.reread < 0 && pa == 0);
    if (REP && pa) {
        const uint64_t sl = (uint64_t)((xcd_id() + c.rot) & (kXcds - 1));
        nlo = N * sl / kXcds;
        nrange = N * (sl + 1) / kXcds - nlo;
    }
    for (int64_t e = s0; e < e0; ++e) {
        while (e >= rend) {
            flush(rp);
            ++r;
            rend = row_ptr[r + 1];
            rp = pos(r);
            fetch(rp, base);
#pragma unroll
            for (int k = 0; k < DIM; ++k) acc[k] = 0.f;
        }
        const int64_t id = rp;  // positions from here on
        for (uint32_t j0 = 0; j0 < total; j0 += PF) {
            int64_t oth[PF];
            float yo[PF][DIM];
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                const uint32_t j = j0 + p;
                oth[p] = id;
                if (j < total) {
                    if (j == 0) {
                        oth[p] = pos(col[e]);
                    } else {
                        int64_t o;
                        do {
                            o = (int64_t)(nlo + uniform_index(g, nrange));
                        } while (o == id);
                        oth[p] = o;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < PF; ++p) fetch(oth[p], yo[p]);
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                if (j0 + p < total) {
                    float yi[DIM], dx[DIM];
#pragma unroll
                    for (int k = 0; k < DIM; ++k) yi[k] = __fadd_rn(base[k], acc[k]);
                    sgd_math<DIM>(yi, yo[p], (j0 + p) == 0, c, Qc, dx);
#pragma unroll
                    for (int k = 0; k < DIM; ++k) acc[k] = __fadd_rn(acc[k], dx[k]);
                    if constexpr (REP) {
                        float cur[DIM];
                        if (reread) {
                            rep_load<DIM>(Y + oth[p] * DIM, cur);
                        } else {
#pragma unroll
                            for (int k = 0; k < DIM; ++k) cur[k] = yo[p][k];
                        }
#pragma unroll
                        for (int k = 0; k < DIM; ++k) cur[k] = __fsub_rn(cur[k], dx[k]);
                        rep_store<DIM>(Y + oth[p] * DIM, cur);
                    } else if (plain_far) {
#pragma unroll
                        for (int k = 0; k < DIM; ++k) Y[oth[p] * DIM + k] = __fsub_rn(yo[p][k], dx[k]);
                    } else if constexpr (DIM % 2 == 0) {
                        // one 64-bit compare-and-swap per coordinate pair: the same float
                        // subtraction as two float atomics, at half the atomic operations; the
                        // prefetched row is the expected value, a concurrent update retries
#pragma unroll
                        for (int k = 0; k < DIM; k += 2) {
                            unsigned long long *addr = (unsigned long long *)&Y[oth[p] * DIM + k];
                            unsigned long long expect =
                                __builtin_bit_cast(unsigned long long, make_float2(yo[p][k], yo[p][k + 1]));
                            while (true) {
                                float2 v = __builtin_bit_cast(float2, expect);
                                float2 nv = make_float2(__fsub_rn(v.x, dx[k]), __fsub_rn(v.y, dx[k + 1]));
                                unsigned long long prev =
                                    atomicCAS(addr, expect, __builtin_bit_cast(unsigned long long, nv));
                                if (prev == expect) break;
                                expect = prev;
                            }
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < DIM; ++k)
                            __hip_atomic_fetch_add(&Y[oth[p] * DIM + k], -dx[k], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        if constexpr (REP) {
            if (c.edge_flush) {
                // the source row's update goes out with every edge (a lost write loses one edge's
                // samples, not the row's whole run) and the chain picks up the others' updates of it
                bool any = false;
#pragma unroll
                for (int k = 0; k < DIM; ++k) any |= acc[k] != 0.f;
                if (any) {
                    float cur[DIM];
                    rep_load<DIM>(Y + rp * DIM, cur);
#pragma unroll
                    for (int k = 0; k < DIM; ++k) {
                        cur[k] = __fadd_rn(cur[k], acc[k]);
                        base[k] = cur[k];
                        acc[k] = 0.f;
                    }
                    rep_store<DIM>(Y + rp * DIM, cur);
                }
            }
        }
    }
    flush(rp);
}


// Sub-epoch `sub` of `S` of every chain of every data set: thread = chain slot (slots past a
// data set's T exit); chain t takes part sub of its static edge chunk, so S consecutive launches
// process the chunk in the reference's order (pcg64 state, Q copy and current row carried over).
// DIM 1..4: d at compile time (modes 2 and 3 available); DIM 0: any d (modes 0 and 1).
template <int DIM, int PF>
__global__ __launch_bounds__(256) void k_sgd_epoch_seg(const FitSeg *__restrict__ s, int B, int64_t slots,
                                                       SampleConst c, int d_rt, int plain_far, int sub, int S) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= slots) return;
    const FitSeg &g = s[seg_find<1>(s, B, gt)];
    const int64_t t = gt - g.chain_off;
    const int64_t T = g.T;
    if (t >= T) return;
    int64_t c0, c1;
    static_chunk(g.E, T, t, c0, c1);
    const int64_t len = c1 - c0;
    const int64_t s0 = c0 + len * sub / S, e0 = c0 + len * (sub + 1) / S;
    float Qc = sub == 0 ? g.qdev[0] : g.qpart[t];
    U128 rng = g.rng[t];
    int64_t r = sub == 0 ? g.chain_row[t] : g.sub_row[t];
    const uint64_t N = (uint64_t)g.N;
    const int mode = g.mode;
    SampleConst ce = c;
    ce.fast_math = 0;  // the exact modes keep the reference's rounding whatever the knob says
    if constexpr (DIM >= 1 && DIM <= 4) {
        if (mode == 3)
            chain_hw<DIM, PF, true>(g.Yrep + (int64_t)xcd_id() * g.N * DIM, g.row_ptr, g.col, s0, e0, r, rng, Qc, c,
                                    N, 0, g.perm_a);
        else if (mode == 2)
            chain_hw<DIM, PF, false>(g.Y, g.row_ptr, g.col, s0, e0, r, rng, Qc, c, N, plain_far);
        else
            chain_exact<DIM>(g.Y, DIM, g.row_ptr, g.col, s0, e0, r, rng, Qc, ce, N, mode);
    } else {
        chain_exact<0>(g.Y, d_rt, g.row_ptr, g.col, s0, e0, r, rng, Qc, ce, N, mode != 0);
    }
    g.rng[t] = rng;
    g.qpart[t] = Qc;
    if (S > 1) g.sub_row[t] = r;  // S > 1 only with sub_row (seg_optimize)
}

// Per-XCD copies of Y (mode 3 data sets): init = 1 copies Y into every copy; otherwise the copies
// are merged, Y + sum over XCDs of (copy - Y) in a fixed order (every update of every copy kept
// once), and the result is written to Y and to every copy. One thread per row.
// With permuted copies (perm_a != 0) thread i handles position i: the base of the merges is the
// permuted Y kept after the copies (Yrep + kXcds N d); init = 1 gathers it from Y (row
// perm_ai i mod N), and k_rep_final scatters it back after the last epoch.
template <int DIM>
__global__ __launch_bounds__(256) void k_rep_sync(const FitSeg *__restrict__ s, int B, int64_t rows, int init) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= rows) return;
    const FitSeg &g = s[seg_find<0>(s, B, gi)];
    if (g.mode != 3) return;
    const int64_t i = gi - g.row_off;
    const int64_t stride = g.N * DIM;
    float *base = g.perm_a ? g.Yrep + kXcds * stride : g.Y;
    float y[DIM];
    if (g.perm_a && init) {
        const int64_t row = (int64_t)(((uint64_t)i * g.perm_ai) % (uint64_t)g.N);
#pragma unroll
        for (int k = 0; k < DIM; ++k) y[k] = g.Y[row * DIM + k];
#pragma unroll
        for (int k = 0; k < DIM; ++k) base[i * DIM + k] = y[k];
    } else {
#pragma unroll
        for (int k = 0; k < DIM; ++k) y[k] = base[i * DIM + k];
    }
    if (!init) {
        float v[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) v[k] = y[k];
#pragma unroll
        for (int x = 0; x < kXcds; ++x) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) v[k] = __fadd_rn(v[k], __fsub_rn(g.Yrep[x * stride + i * DIM + k], y[k]));
        }
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            y[k] = v[k];
            base[i * DIM + k] = v[k];
        }
    }
#pragma unroll
    for (int x = 0; x < kXcds; ++x) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) g.Yrep[x * stride + i * DIM + k] = y[k];
    }
}

// Permuted copies: Y[row] = the merged base at the row's position, after the last merge.
template <int DIM>
__global__ __launch_bounds__(256) void k_rep_final(const FitSeg *__restrict__ s, int B, int64_t rows) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= rows) return;
    const FitSeg &g = s[seg_find<0>(s, B, gi)];
    if (g.mode != 3 || !g.perm_a) return;
    const int64_t i = gi - g.row_off;
    const float *base = g.Yrep + kXcds * g.N * DIM;
    const int64_t row = (int64_t)(((uint64_t)i * g.perm_ai) % (uint64_t)g.N);
#pragma unroll
    for (int k = 0; k < DIM; ++k) g.Y[row * DIM + k] = base[i * DIM + k];
}

// One 1024-thread block per data set; each thread keeps four loads in flight so 2^17 chain
// copies are read at bandwidth instead of one dependent load per thread per step.
__global__ __launch_bounds__(1024) void k_q_reduce_seg(const FitSeg *__restrict__ s, int epoch) {
    __shared__ double red[1024 / 64];
    const FitSeg &g = s[blockIdx.x];
    const int64_t T = g.T;
    const float *__restrict__ qpart = g.qpart;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t t = threadIdx.x;
    for (; t + 3 * 1024 < T; t += 4 * 1024) {
        s0 += (double)qpart[t];
        s1 += (double)qpart[t + 1024];
        s2 += (double)qpart[t + 2 * 1024];
        s3 += (double)qpart[t + 3 * 1024];
    }
    for (; t < T; t += 1024) s0 += (double)qpart[t];
    double sum = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < 1024 / 64; ++w) tot += red[w];
        const float q = (T == 1) ? qpart[0] : (float)(tot / (double)T);
        g.qdev[0] = q;
        g.qdev[1 + epoch] = q;
    }
}

__global__ void k_seg_q0(const FitSeg *__restrict__ s, int B, float Q0) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) s[b].qdev[0] = Q0;
}

template <int DIM>
static void launch_epoch_seg(int PF, dim3 grid, const FitSeg *segs, int B, int64_t slots, const SampleConst &c,
                             int d, int plain_far, int sub, int S, hipStream_t st) {
    switch (PF) {
        case 2: k_sgd_epoch_seg<DIM, 2><<<grid, 256, 0, st>>>(segs, B, slots, c, d, plain_far, sub, S); break;
        case 4: k_sgd_epoch_seg<DIM, 4><<<grid, 256, 0, st>>>(segs, B, slots, c, d, plain_far, sub, S); break;
        case 8: k_sgd_epoch_seg<DIM, 8><<<grid, 256, 0, st>>>(segs, B, slots, c, d, plain_far, sub, S); break;
        default: k_sgd_epoch_seg<DIM, 1><<<grid, 256, 0, st>>>(segs, B, slots, c, d, plain_far, sub, S); break;
    }
}

static void rep_sync(const FitSeg *segs, int B, int64_t rows, int d, int init, hipStream_t st) {
    const int nb = (int)cdiv(rows, 256);
    switch (d) {
        case 1: k_rep_sync<1><<<nb, 256, 0, st>>>(segs, B, rows, init); break;
        case 2: k_rep_sync<2><<<nb, 256, 0, st>>>(segs, B, rows, init); break;
        case 3: k_rep_sync<3><<<nb, 256, 0, st>>>(segs, B, rows, init); break;
        case 4: k_rep_sync<4><<<nb, 256, 0, st>>>(segs, B, rows, init); break;
        default: return;  // mode 3 needs d <= 4
    }
    NCVIS_LAUNCH_CHECK();
}

static void rep_final(const FitSeg *segs, int B, int64_t rows, int d, hipStream_t st) {
    const int nb = (int)cdiv(rows, 256);
    switch (d) {
        case 1: k_rep_final<1><<<nb, 256, 0, st>>>(segs, B, rows); break;
        case 2: k_rep_final<2><<<nb, 256, 0, st>>>(segs, B, rows); break;
        case 3: k_rep_final<3><<<nb, 256, 0, st>>>(segs, B, rows); break;
        case 4: k_rep_final<4><<<nb, 256, 0, st>>>(segs, B, rows); break;
        default: return;
    }
    NCVIS_LAUNCH_CHECK();
}

int sgd_sub_epochs(const SegTotals &tot) { return tot.sub_rows ? std::max(1, tuning().sgd_sync) : 1; }

void seg_optimize(const FitSeg *d_segs, int B, const SegTotals &tot, int d, float Q0, const SgdParams &p,
                  hipStream_t st, hipEvent_t *epoch_events) {
    if (tot.slots <= 0) return;
    const Tuning &tu = tuning();
    const bool rep = tot.replicas && d >= 1 && d <= 4;
    const int S = sgd_sub_epochs(tot);
    k_seg_q0<<<cdiv(B, 64), 64, 0, st>>>(d_segs, B, Q0);
    k_chain_rows_seg<<<cdiv(tot.slots, 256), 256, 0, st>>>(d_segs, B, tot.slots, p.seed);
    NCVIS_LAUNCH_CHECK();
    if (rep) rep_sync(d_segs, B, tot.rows, d, 1, st);
    const dim3 grid((unsigned)cdiv(tot.slots, 256));
    for (int epoch = 0; epoch < p.n_epochs; ++epoch) {
        SampleConst c;
        c.a = p.a;
        c.b = p.b;
        c.alpha_Q = p.alpha_Q;
        const float fe = (float)epoch / (float)p.n_epochs;
        c.step = p.alpha * (1.f - fe * fe);
        c.noise = (uint32_t)p.n_noise[epoch];
        c.noise_f = (float)p.n_noise[epoch];
        c.fast_math = tu.sgd_fast_math;
        c.edge_flush = tu.sgd_edge_flush;
        c.reread = tu.sgd_reread;
        for (int sub = 0; sub < S; ++sub) {
            c.rot = (uint32_t)(3 * epoch + sub);
            hipEvent_t *ev = epoch_events ? epoch_events + 2 * ((size_t)epoch * S + sub) : nullptr;
            if (ev) record_event(ev[0], st);
            switch (d) {
                case 1: launch_epoch_seg<1>(tu.sgd_prefetch, grid, d_segs, B, tot.slots, c, d, tu.sgd_plain_far, sub, S, st); break;
                case 2: launch_epoch_seg<2>(tu.sgd_prefetch, grid, d_segs, B, tot.slots, c, d, tu.sgd_plain_far, sub, S, st); break;
                case 3: launch_epoch_seg<3>(tu.sgd_prefetch, grid, d_segs, B, tot.slots, c, d, tu.sgd_plain_far, sub, S, st); break;
                case 4: launch_epoch_seg<4>(tu.sgd_prefetch, grid, d_segs, B, tot.slots, c, d, tu.sgd_plain_far, sub, S, st); break;
                default:
                    k_sgd_epoch_seg<0, 1><<<grid, 256, 0, st>>>(d_segs, B, tot.slots, c, d, tu.sgd_plain_far, sub, S);
                    break;
            }
            NCVIS_LAUNCH_CHECK();
            if (ev) record_event(ev[1], st);
            if (rep) rep_sync(d_segs, B, tot.rows, d, 0, st);
        }
        k_q_reduce_seg<<<B, 1024, 0, st>>>(d_segs, epoch);
        NCVIS_LAUNCH_CHECK();
    }
    if (rep && tot.permuted) rep_final(d_segs, B, tot.rows, d, st);
}

// One data set with host-known E and chain count (the stage entry points).
void optimize_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t E, float *dY, float *Q,
                  const SgdParams &p, hipStream_t st, float *q_trace, double *epoch_kernel_ms,
                  const SgdWorkspace *ws) {
    if (N < 2 || d < 1) throw HipError("[ncvis_amd] optimize: invalid sizes");
    const int64_t T = p.chains < 1 ? 1 : p.chains;
    DevBuf<int64_t> own_chain_row;
    DevBuf<U128> own_rng;
    DevBuf<float> own_qpart, own_qdev;
    DevBuf<FitSeg> own_seg;
    DevBuf<float> own_rep;
    DevBuf<int64_t> own_sub;
    // atomic_updates 2: the many-chain epochs on per-XCD copies of Y (memory of their own)
    const bool rep = p.atomic_updates == 2;
    if (rep) {
        if (d > 4) throw std::invalid_argument("[ncvis_amd] atomic_updates 2 (per-XCD copies of Y) needs d <= 4");
        if (ws) throw std::invalid_argument("[ncvis_amd] atomic_updates 2 needs device memory for the copies of Y");
        own_rep.alloc((size_t)kXcds * N * d);
    }
    SgdWorkspace w;
    if (ws) {
        w = *ws;
    } else {
        own_chain_row.alloc((size_t)T);
        own_rng.alloc((size_t)T);
        own_qpart.alloc((size_t)T);
        own_qdev.alloc(1 + (size_t)std::max(p.n_epochs, 1));
        own_seg.alloc(1);
        own_sub.alloc((size_t)T);
        w = SgdWorkspace{own_chain_row.p, own_rng.p, own_qpart.p, own_qdev.p, own_seg.p, own_sub.p};
    }
    FitSeg v{};
    v.Y = dY;
    v.row_ptr = d_row_ptr;
    v.col = d_col;
    v.N = N;
    v.T_cap = T;
    v.chain_row = w.chain_row;
    v.rng = w.rng;
    v.qpart = w.qpart;
    v.qdev = w.qdev;
    v.pairs_ok = ((uintptr_t)dY % 8 == 0) ? 1 : 0;
    v.E = E;
    v.T = T;
    v.mode = rep ? 3 : (p.atomic_updates ? ((d <= 4 && (d % 2 != 0 || v.pairs_ok)) ? 2 : 1) : 0);
    v.Yrep = rep ? own_rep.p : nullptr;
    v.sub_row = w.sub_row;
    seg_set(w.seg, v, st);
    SegTotals tot;
    tot.rows = N;
    tot.slots = T;
    tot.replicas = rep;
    tot.sub_rows = w.sub_row != nullptr;
    std::vector<hipEvent_t> ev;
    struct EvGuard {
        std::vector<hipEvent_t> &e;
        ~EvGuard() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } evg{ev};
    const int launches = std::max(p.n_epochs, 0) * sgd_sub_epochs(tot);
    if (epoch_kernel_ms) {
        ev.resize(2 * (size_t)launches);
        for (auto &x : ev) {
            x = nullptr;
            NCVIS_HIP_CHECK(hipEventCreate(&x));
        }
    }
    seg_optimize(w.seg, 1, tot, d, *Q, p, st, epoch_kernel_ms ? ev.data() : nullptr);
    NCVIS_HIP_CHECK(hipMemcpyAsync(Q, w.qdev, sizeof(float), hipMemcpyDeviceToHost, st));
    if (q_trace && p.n_epochs > 0)
        NCVIS_HIP_CHECK(
            hipMemcpyAsync(q_trace, w.qdev + 1, sizeof(float) * p.n_epochs, hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (epoch_kernel_ms) *epoch_kernel_ms = epoch_time_ms(ev.data(), launches);
}

double epoch_time_ms(const hipEvent_t *ev, int n_launches) {
    double tot = 0.0;
    for (int l = 0; l < n_launches; ++l) {
        float ms = 0.f;
        NCVIS_HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * l], ev[2 * l + 1]));
        tot += ms;
    }
    return tot;
}

// ------------------------------------------------------------------------ graph checks
// err bit 1: row_ptr[0] != 0 or row_ptr[N] != E; bit 2: a row end below its start; bit 4: a
// column index outside [0, N); bit 8: a negative index outside [0, N).
__global__ void k_check_graph(const int64_t *__restrict__ row_ptr, int64_t N, const int32_t *__restrict__ col,
                              int64_t E, const int32_t *__restrict__ neg, int64_t n_neg, int *__restrict__ err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0 && (row_ptr[0] != 0 || row_ptr[N] != E)) atomicOr(err, 1);
    for (int64_t i = t0; i < N; i += stride)
        if (row_ptr[i + 1] < row_ptr[i]) atomicOr(err, 2);
    for (int64_t e = t0; e < E; e += stride)
        if (col[e] < 0 || col[e] >= N) atomicOr(err, 4);
    for (int64_t e = t0; e < n_neg; e += stride)
        if (neg[e] < 0 || neg[e] >= N) atomicOr(err, 8);
}

void check_graph(const int64_t *d_row_ptr, int64_t N, const int32_t *d_col, int64_t E, const int32_t *d_neg,
                 int64_t n_neg, hipStream_t st, int *d_err) {
    if (E > 0 && !d_col) throw std::invalid_argument("[ncvis_amd] edges given without a column array");
    DevBuf<int> own;
    if (!d_err) {
        own.alloc(1);
        d_err = own.p;
    }
    NCVIS_HIP_CHECK(hipMemsetAsync(d_err, 0, sizeof(int), st));
    const int64_t work = std::max<int64_t>(std::max<int64_t>(N, E), n_neg);
    const int nb = (int)std::min<int64_t>(4096, std::max<int64_t>(1, cdiv(work, 256)));
    k_check_graph<<<nb, 256, 0, st>>>(d_row_ptr, N, d_col, E, d_neg, n_neg, d_err);
    NCVIS_LAUNCH_CHECK();
    int herr = 0;
    NCVIS_HIP_CHECK(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (herr & 1) throw std::invalid_argument("[ncvis_amd] row_ptr must start at 0 and end at the edge count");
    if (herr & 2) throw std::invalid_argument("[ncvis_amd] row_ptr must be non-decreasing");
    if (herr & 4) throw std::invalid_argument("[ncvis_amd] column index outside [0, N)");
    if (herr & 8) throw std::invalid_argument("[ncvis_amd] negative-sample index outside [0, N)");
}

// ------------------------------------------------------------------------ NCE loss
__global__ void k_edge_src(const int64_t *__restrict__ row_ptr, int64_t N, int32_t *__restrict__ src) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) src[e] = (int32_t)i;
}

// s = log Ph - Q - log n, Ph = 1/(1 + a d2^b); positive term softplus(-s), negative softplus(s).
// dL/ds: positive -(1 - sigmoid(s)), negative sigmoid(s); ds/dd2 = -a b d2^(b-1) Ph.
__global__ __launch_bounds__(256) void k_nce_loss(const float *__restrict__ Y, int d, const int32_t *__restrict__ src,
                                                  const int32_t *__restrict__ col, int64_t E,
                                                  const int32_t *__restrict__ neg, int n, float Q, float a, float b,
                                                  double *__restrict__ partial, float *__restrict__ gY) {
    __shared__ double red_l[256], red_q[256];
    double L = 0.0, gQ = 0.0;
    const float logn = logf((float)n);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = src[e];
        for (int m = 0; m <= n; ++m) {
            const int64_t o = (m == 0) ? col[e] : neg[e * n + (m - 1)];
            float d2 = 0.f;
            for (int c = 0; c < d; ++c) {
                float t = Y[i * d + c] - Y[o * d + c];
                d2 = fmaf(t, t, d2);
            }
            float pw = powf(d2, b);
            float logPh = -log1pf(a * pw);
            float s = logPh - Q - logn;
            float sig = 1.f / (1.f + expf(-s));
            float g;
            if (m == 0) {
                L += (double)((s > 0.f) ? log1pf(expf(-s)) : (-s + log1pf(expf(s))));
                g = -(1.f - sig);
            } else {
                L += (double)((s > 0.f) ? (s + log1pf(expf(-s))) : log1pf(expf(s)));
                g = sig;
            }
            gQ -= (double)g;
            if (gY && d2 > 0.f) {
                float dsdd2 = -a * b * (pw / d2) / (1.f + a * pw);
                float coef = 2.f * g * dsdd2;
                for (int c = 0; c < d; ++c) {
                    float gi = coef * (Y[i * d + c] - Y[o * d + c]);
                    atomicAdd(&gY[i * d + c], gi);
                    atomicAdd(&gY[o * d + c], -gi);
                }
            }
        }
    }
    red_l[threadIdx.x] = L;
    red_q[threadIdx.x] = gQ;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red_l[threadIdx.x] += red_l[threadIdx.x + w];
            red_q[threadIdx.x] += red_q[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red_l[0];
        partial[2 * blockIdx.x + 1] = red_q[0];
    }
}

__global__ void k_nce_finalize(const double *__restrict__ partial, int nb, double *__restrict__ loss,
                               double *__restrict__ gQ) {
    if (threadIdx.x != 0) return;
    double l = 0, q = 0;
    for (int b = 0; b < nb; ++b) {
        l += partial[2 * b];
        q += partial[2 * b + 1];
    }
    *loss = l;
    if (gQ) *gQ = q;
}

void nce_loss_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t E, const int32_t *d_neg,
                  int n, const float *dY_in, float Q, float a, float b, double *d_loss, float *d_gradY,
                  double *d_gradQ, hipStream_t st) {
    if (N < 1 || d < 1 || n < 1) throw HipError("[ncvis_amd] nce_loss: needs N >= 1, d >= 1, n >= 1");
    DevBuf<int32_t> src((size_t)std::max<int64_t>(E, 1));
    k_edge_src<<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, N, src.p);
    NCVIS_LAUNCH_CHECK();
    const int nb = (int)std::min<int64_t>(2048, std::max<int64_t>(1, cdiv(E, 256)));
    DevBuf<double> partial((size_t)2 * nb);
    k_nce_loss<<<nb, 256, 0, st>>>(dY_in, d, src.p, d_col, E, d_neg, n, Q, a, b, partial.p, d_gradY);
    NCVIS_LAUNCH_CHECK();
    k_nce_finalize<<<1, 64, 0, st>>>(partial.p, nb, d_loss, d_gradQ);
    NCVIS_LAUNCH_CHECK();
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace ncvis_hip
