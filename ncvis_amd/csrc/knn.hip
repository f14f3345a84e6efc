// Exact all-pairs kNN for the NCVis fit path on gfx950.
//
// Replaces buildKNN + findKNN (src/ncvis.cpp:72-153, hnswlib HNSW) with an exact search that
// keeps findKNN's output contract: per row, the k+1 smallest (distance, label) pairs, the single
// smallest dropped (normally the point itself), the remaining k stored farthest first
// (src/ncvis.cpp:142-147). Distances follow hnswlib's spaces: squared L2 for euclidean,
// 1 - <x, y> for inner_product / cosine / correlation, on rows preprocessed as in
// src/ncvis.cpp:32-70.
//
// Pipeline:
//   k_preprocess  - cosine/correlation row preprocessing (one thread per row, FP32)
//   k_pack        - candidate matrix in the GEMM precision (FP32, BF16 or FP16), zero padded to
//                   Dpad = 32*DK columns, plus per-row score offsets cn (||x||^2 for L2, 0 for IP,
//                   +inf for padding rows)
//   k_knn_tiles   - MFMA tiles: a workgroup owns 4 waves x 32*QB queries whose B fragments live
//                   in registers; candidate rows stream through double-buffered LDS stages; every
//                   32x32 MFMA block is scored (cn + beta * <q, c>) and filtered into per-lane
//                   sorted top-KL lists (two half-lists per query, one per lane half)
//   k_knn_rerank  - exact FP32 distances for the 2*KL candidates of each query, rank by
//                   (distance, label), drop rank 0, write ranks k..1 (farthest first)
//   k_knn_generic - the remaining shapes: exact FP32 scan (LDS-tiled) with a max-heap per query
#include <cfloat>
#include <climits>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// ------------------------------------------------------------------------ preprocess
// src/ncvis.cpp:32-70; sums are sequential and uncontracted like the reference's loops.
__global__ void k_preprocess(const float *__restrict__ X, int64_t N, int D, int dist, float *__restrict__ Xp) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *x = X + i * D;
    float *o = Xp + i * D;
    if (dist == DIST_CORRELATION) {
        float M = 0.f;
        for (int j = 0; j < D; ++j) M = __fadd_rn(M, x[j]);
        M = M / (float)D;
        for (int j = 0; j < D; ++j) o[j] = __fsub_rn(x[j], M);
    } else {
        for (int j = 0; j < D; ++j) o[j] = x[j];
    }
    if (dist == DIST_CORRELATION || dist == DIST_COSINE) {
        float s = 0.f;
        for (int j = 0; j < D; ++j) s = __fadd_rn(s, __fmul_rn(o[j], o[j]));
        if (s != 0.f) {
            float r = sqrtf(s);
            for (int j = 0; j < D; ++j) o[j] = o[j] / r;
        }
    }
}

// ------------------------------------------------------------------------ packing
__device__ inline uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Element k of a padded row as stored for MFMA. 16-bit precisions keep the natural order
// (lane half h of a 32x32x16 step reads k = 16 s + 8 h .. +7, contiguous). FP32 (32x32x2:
// lane half h reads k = 2 s + h) is stored permuted, position h*(Dpad/2) + s, so a lane's
// operands are contiguous too. A and B use the same permutation, so dot products are unchanged.
template <int P>
__global__ void k_pack(const float *__restrict__ Xp, int64_t N, int64_t Npad, int D, int Dpad, int is_l2,
                       void *__restrict__ Xh, float *__restrict__ cn) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Npad) return;
    float s = 0.f;
    if (P == PREC_FP32) {
        float *o = (float *)Xh + i * Dpad;
        const int half = Dpad / 2;
        for (int p = 0; p < Dpad; ++p) {
            int h = p / half, st = p - h * half;
            int kk = 2 * st + h;
            float v = (i < N && kk < D) ? Xp[i * D + kk] : 0.f;
            o[p] = v;
            s = __fadd_rn(s, __fmul_rn(v, v));
        }
    } else {
        uint16_t *o = (uint16_t *)Xh + i * Dpad;
        for (int kk = 0; kk < Dpad; ++kk) {
            float v = (i < N && kk < D) ? Xp[i * D + kk] : 0.f;
            float r;
            if (P == PREC_BF16) {
                uint16_t b = f32_to_bf16_rne(v);
                o[kk] = b;
                r = bf16_to_f32(b);
            } else {
                _Float16 hv = (_Float16)v;
                o[kk] = __builtin_bit_cast(uint16_t, hv);
                r = (float)hv;
            }
            s = __fadd_rn(s, __fmul_rn(r, r));
        }
    }
    cn[i] = (i < N) ? (is_l2 ? s : 0.f) : INFINITY;
}

// ------------------------------------------------------------------------ top-k lists
// Sorted ascending list: x goes to the first slot p with x < ld[p], the tail shifts down one slot
// and the last entry drops out. Among equal scores the earlier entry stays first and the later
// one drops first, so with candidates visited in ascending label order a list holds the KL
// smallest (score, label) pairs. Every new slot is a function of the old list only: the score
// of slot p is med3(x, ld[p-1], ld[p]) (the list is sorted), its label comes from slot p-1, x or
// itself, so the update runs in place without swaps. Requires x < ld[KL-1] (x not NaN).
template <int KL>
__device__ __forceinline__ void list_insert(float (&ld)[KL], int (&li)[KL], float x, int xi) {
    bool c[KL];
#pragma unroll
    for (int p = 0; p < KL; ++p) c[p] = x < ld[p];
#pragma unroll
    for (int p = KL - 1; p >= 1; --p) {
        ld[p] = __builtin_amdgcn_fmed3f(x, ld[p - 1], ld[p]);
        li[p] = c[p - 1] ? li[p - 1] : (c[p] ? xi : li[p]);
    }
    ld[0] = c[0] ? x : ld[0];
    li[0] = c[0] ? xi : li[0];
}

template <int P>
struct Frag;
template <>
struct Frag<PREC_FP32> {
    using T = float;  // one k element per lane per 32x32x2 step
    static constexpr int ESZ = 4;
};
template <>
struct Frag<PREC_BF16> {
    using T = uint4;  // eight bf16 per lane per 32x32x16 step
    static constexpr int ESZ = 2;
};
template <>
struct Frag<PREC_FP16> {
    using T = uint4;
    static constexpr int ESZ = 2;
};

template <int P>
__device__ __forceinline__ f32x16_t mfma_step(typename Frag<P>::T a, typename Frag<P>::T b, f32x16_t c) {
    if constexpr (P == PREC_FP32) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    } else if constexpr (P == PREC_BF16) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                       __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a),
                                                      __builtin_bit_cast(f16x8_t, b), c, 0, 0, 0);
    }
}

// MFMA 32x32 output layout: lane l holds column l%32 and rows 8*(i/4) + 4*(l/32) + (i%4),
// i = 0..15. Candidates are the A rows, queries the B columns, so a lane holds one query.
template <int P, int QB, int KL, int DK, int SM = 1>
__global__ __launch_bounds__(256) void k_knn_tiles(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                   int64_t N, int64_t Npad, float beta,
                                                   int32_t *__restrict__ cand) {
    using FT = typename Frag<P>::T;
    constexpr int ESZ = Frag<P>::ESZ;
    constexpr int DP = 32 * DK;                       // padded row length (elements)
    constexpr int KS = (P == PREC_FP32) ? DP / 2 : DP / 16;  // MFMA k-steps per row
    constexpr int ROWB = DP * ESZ;                    // row bytes in HBM
    constexpr int CB = ((ROWB >= 512) ? 32 : 64) * SM;  // candidate rows per LDS stage
    constexpr int LROWB = ROWB + 16;                  // padded LDS row (bank spread)
    constexpr int NV = CB * ROWB / 16 / 256;          // uint4 per thread per stage
    constexpr int CHUNKS = ROWB / 16;                 // uint4 per row
    static_assert(NV >= 1 && NV * 256 * 16 == CB * ROWB, "stage must split evenly over 256 threads");

    __shared__ __align__(16) unsigned char lds[2][CB * LROWB];
    __shared__ __align__(16) float ldc[2][CB];  // the stage's score offsets travel with its rows

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int col = lane & 31;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * QB);

    // query (B) fragments, resident for the whole candidate sweep
    FT qf[QB][KS];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        const unsigned char *row = Xh + (q0 + qb * 32 + col) * (int64_t)ROWB;
        if constexpr (P == PREC_FP32) {
#pragma unroll
            for (int s = 0; s < KS; s += 4) {
                float4 v = *(const float4 *)(row + (h * (DP / 2) + s) * 4);
                qf[qb][s] = v.x;
                qf[qb][s + 1] = v.y;
                qf[qb][s + 2] = v.z;
                qf[qb][s + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) qf[qb][s] = *(const uint4 *)(row + s * 32 + h * 16);
        }
    }

    float ld[QB][KL];
    int li[QB][KL];
    float thr[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
#pragma unroll
        for (int p = 0; p < KL; ++p) {
            ld[qb][p] = INFINITY;
            li[qb][p] = -1;
        }
        thr[qb] = INFINITY;
    }

    const int64_t nstages = Npad / CB;
    const bool cn_lane = threadIdx.x < CB / 4;  // these threads carry the stage's cn (float4 each)
    uint4 stage[NV];
    float4 stage_cn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(Xh + (int64_t)(v * 256 + threadIdx.x) * 16);
    if (cn_lane) stage_cn = *(const float4 *)(cn + 4 * threadIdx.x);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        int u = v * 256 + threadIdx.x;
        int r = u / CHUNKS, c = u - r * CHUNKS;
        *(uint4 *)(&lds[0][r * LROWB + c * 16]) = stage[v];
    }
    if (cn_lane) *(float4 *)(&ldc[0][4 * threadIdx.x]) = stage_cn;
    __syncthreads();

    for (int64_t st = 0; st < nstages; ++st) {
        const int buf = (int)(st & 1);
        {
            // next stage (the last iteration reloads its own stage; the copy is never read)
            const int64_t nx = (st + 1 < nstages) ? st + 1 : st;
            const unsigned char *nb = Xh + nx * CB * (int64_t)ROWB;
#pragma unroll
            for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(nb + (int64_t)(v * 256 + threadIdx.x) * 16);
            if (cn_lane) stage_cn = *(const float4 *)(cn + nx * CB + 4 * threadIdx.x);
        }
#pragma unroll
        for (int rb = 0; rb < CB / 32; ++rb) {
            const int64_t c0 = st * CB + rb * 32;
            const unsigned char *arow = &lds[buf][(rb * 32 + col) * LROWB];
            f32x16_t acc[QB];
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) acc[qb] = f32x16_t{};
            if constexpr (P == PREC_FP32) {
#pragma unroll
                for (int s = 0; s < KS; s += 4) {
                    float4 a4 = *(const float4 *)(arow + (h * (DP / 2) + s) * 4);
#pragma unroll
                    for (int qb = 0; qb < QB; ++qb) {
                        acc[qb] = mfma_step<P>(a4.x, qf[qb][s], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.y, qf[qb][s + 1], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.z, qf[qb][s + 2], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.w, qf[qb][s + 3], acc[qb]);
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    uint4 a = *(const uint4 *)(arow + s * 32 + h * 16);
#pragma unroll
                    for (int qb = 0; qb < QB; ++qb) acc[qb] = mfma_step<P>(a, qf[qb][s], acc[qb]);
                }
            }
            // score offsets of this lane's 16 candidate rows: c0 + 8g + 4h + j
            float cnv[16];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float4 v = *(const float4 *)(&ldc[buf][rb * 32 + 8 * g + 4 * h]);
                cnv[4 * g] = v.x;
                cnv[4 * g + 1] = v.y;
                cnv[4 * g + 2] = v.z;
                cnv[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) {
                float m = INFINITY;
#pragma unroll
                for (int i = 0; i < 16; ++i) m = fminf(m, fmaf(beta, acc[qb][i], cnv[i]));
                if (__any(m < thr[qb])) {
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        float sc = fmaf(beta, acc[qb][i], cnv[i]);
                        bool pass = sc < thr[qb];
                        if (__any(pass)) {
                            if (pass) {
                                int ci = (int)(c0 + 8 * (i >> 2) + 4 * h + (i & 3));
                                list_insert<KL>(ld[qb], li[qb], sc, ci);
                                thr[qb] = ld[qb][KL - 1];
                            }
                        }
                    }
                }
            }
        }
        {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                int u = v * 256 + threadIdx.x;
                int r = u / CHUNKS, c = u - r * CHUNKS;
                *(uint4 *)(&lds[buf ^ 1][r * LROWB + c * 16]) = stage[v];
            }
            if (cn_lane) *(float4 *)(&ldc[buf ^ 1][4 * threadIdx.x]) = stage_cn;
        }
        __syncthreads();
    }

#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        int64_t q = q0 + qb * 32 + col;
        if (q < N) {
            int32_t *o = cand + q * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = li[qb][p];
        }
    }
}

// ------------------------------------------------------------------------ skewed two-set tiles
// Scores of one 32x32 block (held in the accumulator, overwritten in place by cn + beta * dot)
// filtered into a lane's sorted list; candidates are visited in ascending label order.
template <int KL>
__device__ __forceinline__ void block_insert(float (&ld)[KL], int (&li)[KL], float &thr, const f32x16_t &sc,
                                             int64_t c0, int h) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const bool pass = sc[i] < thr;
        if (__any(pass)) {
            if (pass) {
                list_insert<KL>(ld, li, sc[i], (int)(c0 + 8 * (i >> 2) + 4 * h + (i & 3)));
                thr = ld[KL - 1];
            }
        }
    }
}

// cn + beta * dot for the lane's 16 rows of a block (offsets read from the stage's LDS copy);
// returns the lane's smallest score.
__device__ __forceinline__ float block_scores(f32x16_t &acc, const float *pc, int h, float beta) {
    float m = INFINITY;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 v = *(const float4 *)(pc + 8 * g + 4 * h);
        acc[4 * g] = fmaf(beta, acc[4 * g], v.x);
        acc[4 * g + 1] = fmaf(beta, acc[4 * g + 1], v.y);
        acc[4 * g + 2] = fmaf(beta, acc[4 * g + 2], v.z);
        acc[4 * g + 3] = fmaf(beta, acc[4 * g + 3], v.w);
        m = fminf(m, fminf(fminf(acc[4 * g], acc[4 * g + 1]), fminf(acc[4 * g + 2], acc[4 * g + 3])));
    }
    return m;
}

// Same tiles, lists and visiting order as k_knn_tiles at QB = 2 (16-bit GEMM), with the two query
// sets of a wave skewed by one block: the MFMA chain of set 0 for block b is issued together
// with the scoring of set 1 for block b-1, and the chain of set 1 for block b together with the
// scoring of set 0 for block b, so one wave keeps its MFMA pipe busy while its VALU filters the
// other set. Score offsets are triple-buffered in LDS because set 1 reads the previous stage's
// offsets after the stage barrier. The candidate lists equal k_knn_tiles' bit for bit.
template <int P, int KL, int DK, int SM>
__global__ __launch_bounds__(256) void k_knn_tiles_sk(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                      int64_t N, int64_t Npad, float beta,
                                                      int32_t *__restrict__ cand) {
    static_assert(P != PREC_FP32, "16-bit candidate GEMM only");
    constexpr int QB = 2;
    constexpr int ESZ = 2;
    constexpr int DP = 32 * DK;
    constexpr int KS = DP / 16;
    constexpr int ROWB = DP * ESZ;
    constexpr int CB = ((ROWB >= 512) ? 32 : 64) * SM;
    constexpr int LROWB = ROWB + 16;
    constexpr int NV = CB * ROWB / 16 / 256;
    constexpr int CHUNKS = ROWB / 16;
    constexpr int NB = CB / 32;
    static_assert(NV >= 1 && NV * 256 * 16 == CB * ROWB, "stage must split evenly over 256 threads");

    __shared__ __align__(16) unsigned char lds[2][CB * LROWB];
    __shared__ __align__(16) float ldc[3][CB];
    __shared__ __align__(16) float ldinf[32];

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int col = lane & 31;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * QB);

    uint4 qf[QB][KS];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        const unsigned char *row = Xh + (q0 + qb * 32 + col) * (int64_t)ROWB;
#pragma unroll
        for (int s = 0; s < KS; ++s) qf[qb][s] = *(const uint4 *)(row + s * 32 + h * 16);
    }
    float ld[QB][KL];
    int li[QB][KL];
    float thr[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
#pragma unroll
        for (int p = 0; p < KL; ++p) {
            ld[qb][p] = INFINITY;
            li[qb][p] = -1;
        }
        thr[qb] = INFINITY;
    }

    const int64_t nstages = Npad / CB;
    const bool cn_lane = threadIdx.x < CB / 4;
    uint4 stage[NV];
    float4 stage_cn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(Xh + (int64_t)(v * 256 + threadIdx.x) * 16);
    if (cn_lane) stage_cn = *(const float4 *)(cn + 4 * threadIdx.x);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        int u = v * 256 + threadIdx.x;
        int r = u / CHUNKS, c = u - r * CHUNKS;
        *(uint4 *)(&lds[0][r * LROWB + c * 16]) = stage[v];
    }
    if (cn_lane) *(float4 *)(&ldc[0][4 * threadIdx.x]) = stage_cn;
    if (threadIdx.x < 32) ldinf[threadIdx.x] = INFINITY;
    __syncthreads();

    // set 1's accumulator of the previous block; before the first block it scores +inf
    f32x16_t acc1 = f32x16_t{};
    const float *pc1 = ldinf;
    int64_t pc0 = 0;
    int cb = 0;  // st % 3
    for (int64_t st = 0; st < nstages; ++st) {
        const int buf = (int)(st & 1);
        {
            const int64_t nx = (st + 1 < nstages) ? st + 1 : st;
            const unsigned char *nb = Xh + nx * CB * (int64_t)ROWB;
#pragma unroll
            for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(nb + (int64_t)(v * 256 + threadIdx.x) * 16);
            if (cn_lane) stage_cn = *(const float4 *)(cn + nx * CB + 4 * threadIdx.x);
        }
#pragma unroll
        for (int rb = 0; rb < NB; ++rb) {
            const int64_t c0 = st * CB + rb * 32;
            const unsigned char *arow = &lds[buf][(rb * 32 + col) * LROWB];
            uint4 a[KS];
#pragma unroll
            for (int s = 0; s < KS; ++s) a[s] = *(const uint4 *)(arow + s * 32 + h * 16);
            f32x16_t acc0 = f32x16_t{};
#pragma unroll
            for (int s = 0; s < KS; ++s) acc0 = mfma_step<P>(a[s], qf[0][s], acc0);
            const float m1 = block_scores(acc1, pc1, h, beta);
            if (__any(m1 < thr[1])) block_insert<KL>(ld[1], li[1], thr[1], acc1, pc0, h);
            acc1 = f32x16_t{};
#pragma unroll
            for (int s = 0; s < KS; ++s) acc1 = mfma_step<P>(a[s], qf[1][s], acc1);
            const float *pc = &ldc[cb][rb * 32];
            const float m0 = block_scores(acc0, pc, h, beta);
            if (__any(m0 < thr[0])) block_insert<KL>(ld[0], li[0], thr[0], acc0, c0, h);
            pc1 = pc;
            pc0 = c0;
        }
        const int cbn = (cb == 2) ? 0 : cb + 1;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            int u = v * 256 + threadIdx.x;
            int r = u / CHUNKS, c = u - r * CHUNKS;
            *(uint4 *)(&lds[buf ^ 1][r * LROWB + c * 16]) = stage[v];
        }
        if (cn_lane) *(float4 *)(&ldc[cbn][4 * threadIdx.x]) = stage_cn;
        cb = cbn;
        __syncthreads();
    }
    {
        const float m1 = block_scores(acc1, pc1, h, beta);
        if (__any(m1 < thr[1])) block_insert<KL>(ld[1], li[1], thr[1], acc1, pc0, h);
    }

#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        int64_t q = q0 + qb * 32 + col;
        if (q < N) {
            int32_t *o = cand + q * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = li[qb][p];
        }
    }
}

// ------------------------------------------------------------------------ buffered selection
// Variant of k_knn_tiles whose hot loop holds no list: each (query, lane half) keeps only a
// threshold and appends candidates that beat it, {score bits, label}, to its own buffer of CAP
// entries in HBM (a few hundred appends over a whole sweep). When a buffer could overflow, the
// wave re-selects the best KL of every buffer in lockstep (one sorted-insertion pass per buffered
// slot, registers live only inside that pass) and tightens the thresholds. Entries are appended
// in ascending label order and re-selected in buffer order with strict '<', so the final lists
// equal one-at-a-time insertion.
template <int KL>
__device__ __forceinline__ void buf_reselect(uint2 *__restrict__ b, int &cnt, float &thr) {
    float ld[KL];
    int li[KL];
#pragma unroll
    for (int p = 0; p < KL; ++p) {
        ld[p] = INFINITY;
        li[p] = -1;
    }
    for (int j = 0; __any(j < cnt); ++j) {
        const uint2 e = (j < cnt) ? b[j] : make_uint2(0x7f800000u, 0u);
        list_insert<KL>(ld, li, __uint_as_float(e.x), (int)e.y);
    }
    int nvalid = 0;
#pragma unroll
    for (int p = 0; p < KL; ++p) {
        b[p] = make_uint2(__float_as_uint(ld[p]), (unsigned)li[p]);
        nvalid += (li[p] >= 0) ? 1 : 0;
    }
    cnt = nvalid;
    thr = (nvalid == KL) ? ld[KL - 1] : INFINITY;
}

template <int P, int QB, int KL, int DK>
__global__ __launch_bounds__(256) void k_knn_tiles_buf(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                       int64_t N, int64_t Npad, float beta,
                                                       uint2 *__restrict__ sbuf, int32_t *__restrict__ cand) {
    using FT = typename Frag<P>::T;
    constexpr int ESZ = Frag<P>::ESZ;
    constexpr int DP = 32 * DK;
    constexpr int KS = (P == PREC_FP32) ? DP / 2 : DP / 16;
    constexpr int ROWB = DP * ESZ;
    constexpr int CB = (ROWB >= 512) ? 32 : 64;
    constexpr int LROWB = ROWB + 16;
    constexpr int NV = CB * ROWB / 16 / 256;
    constexpr int CHUNKS = ROWB / 16;
    constexpr int CAP = 4 * KL;
    static_assert(NV >= 1 && NV * 256 * 16 == CB * ROWB, "stage must split evenly over 256 threads");

    __shared__ __align__(16) unsigned char lds[2][CB * LROWB];

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int col = lane & 31;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * QB);

    FT qf[QB][KS];
    uint2 *mb[QB];
    float thr[QB];
    int cnt[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        const int64_t q = q0 + qb * 32 + col;
        const unsigned char *row = Xh + q * (int64_t)ROWB;
        if constexpr (P == PREC_FP32) {
#pragma unroll
            for (int s = 0; s < KS; s += 4) {
                float4 v = *(const float4 *)(row + (h * (DP / 2) + s) * 4);
                qf[qb][s] = v.x;
                qf[qb][s + 1] = v.y;
                qf[qb][s + 2] = v.z;
                qf[qb][s + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) qf[qb][s] = *(const uint4 *)(row + s * 32 + h * 16);
        }
        mb[qb] = sbuf + (q * 2 + h) * (int64_t)CAP;
        thr[qb] = INFINITY;
        cnt[qb] = 0;
    }

    const int64_t nstages = Npad / CB;
    uint4 stage[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(Xh + (int64_t)(v * 256 + threadIdx.x) * 16);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        int u = v * 256 + threadIdx.x;
        int r = u / CHUNKS, c = u - r * CHUNKS;
        *(uint4 *)(&lds[0][r * LROWB + c * 16]) = stage[v];
    }
    __syncthreads();

    for (int64_t st = 0; st < nstages; ++st) {
        const int buf = (int)(st & 1);
        {
            const int64_t nx = (st + 1 < nstages) ? st + 1 : st;
            const unsigned char *nb = Xh + nx * CB * (int64_t)ROWB;
#pragma unroll
            for (int v = 0; v < NV; ++v) stage[v] = *(const uint4 *)(nb + (int64_t)(v * 256 + threadIdx.x) * 16);
        }
#pragma unroll
        for (int rb = 0; rb < CB / 32; ++rb) {
            const int64_t c0 = st * CB + rb * 32;
            const unsigned char *arow = &lds[buf][(rb * 32 + col) * LROWB];
            f32x16_t acc[QB];
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) acc[qb] = f32x16_t{};
            if constexpr (P == PREC_FP32) {
#pragma unroll
                for (int s = 0; s < KS; s += 4) {
                    float4 a4 = *(const float4 *)(arow + (h * (DP / 2) + s) * 4);
#pragma unroll
                    for (int qb = 0; qb < QB; ++qb) {
                        acc[qb] = mfma_step<P>(a4.x, qf[qb][s], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.y, qf[qb][s + 1], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.z, qf[qb][s + 2], acc[qb]);
                        acc[qb] = mfma_step<P>(a4.w, qf[qb][s + 3], acc[qb]);
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    uint4 a = *(const uint4 *)(arow + s * 32 + h * 16);
#pragma unroll
                    for (int qb = 0; qb < QB; ++qb) acc[qb] = mfma_step<P>(a, qf[qb][s], acc[qb]);
                }
            }
            float cnv[16];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float4 v = *(const float4 *)(cn + c0 + 8 * g + 4 * h);
                cnv[4 * g] = v.x;
                cnv[4 * g + 1] = v.y;
                cnv[4 * g + 2] = v.z;
                cnv[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) {
                float m = INFINITY;
#pragma unroll
                for (int i = 0; i < 16; ++i) m = fminf(m, fmaf(beta, acc[qb][i], cnv[i]));
                if (__any(m < thr[qb])) {
                    if (__any(cnt[qb] > CAP - 16)) buf_reselect<KL>(mb[qb], cnt[qb], thr[qb]);
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const float sc = fmaf(beta, acc[qb][i], cnv[i]);
                        if (sc < thr[qb]) {
                            mb[qb][cnt[qb]] = make_uint2(__float_as_uint(sc), (unsigned)(c0 + 8 * (i >> 2) + 4 * h + (i & 3)));
                            ++cnt[qb];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            int u = v * 256 + threadIdx.x;
            int r = u / CHUNKS, c = u - r * CHUNKS;
            *(uint4 *)(&lds[buf ^ 1][r * LROWB + c * 16]) = stage[v];
        }
        __syncthreads();
    }

#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        buf_reselect<KL>(mb[qb], cnt[qb], thr[qb]);
        const int64_t q = q0 + qb * 32 + col;
        if (q < N) {
            int32_t *o = cand + q * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = (int32_t)mb[qb][p].y;
        }
    }
}

// ------------------------------------------------------------------------ LDS-DMA tile kernel
// Same tile and selection as k_knn_tiles, but candidate stages are fetched with
// global_load_lds_dwordx4 straight into LDS (no register staging), three stages in flight. LDS
// rows are unpadded; the 16-byte chunk c of row r sits in slot c ^ (r % CHUNKS), so the
// fragment reads of 32 consecutive rows at one chunk index hit distinct banks. Each wave keeps
// its own copy of a stage's score offsets, fetched by its first CB/4 lanes.
#define NCVIS_AS1(p) ((const void __attribute__((address_space(1))) *)(p))
#define NCVIS_AS3(p) ((void __attribute__((address_space(3))) *)(p))

template <int NV, int CB, int ROWB, int CHUNKS>
__device__ __forceinline__ void dma_stage(const unsigned char *__restrict__ Xh, const float *__restrict__ cn,
                                          int64_t stage, unsigned char *lbuf, float *lcn_wave, int wave, int lane) {
    const unsigned char *base = Xh + stage * (int64_t)(CB * ROWB);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const int p = (v * 4 + wave) * 64 + lane;  // 16-byte LDS slot this lane fills
        const int r = p / CHUNKS, s = p % CHUNKS;
        const int c = s ^ (r & (CHUNKS - 1));
        __builtin_amdgcn_global_load_lds(NCVIS_AS1(base + r * ROWB + c * 16), NCVIS_AS3(lbuf + (v * 4 + wave) * 1024),
                                         16, 0, 0);
    }
    if (lane < CB / 4)
        __builtin_amdgcn_global_load_lds(NCVIS_AS1(cn + stage * CB + 4 * lane), NCVIS_AS3(lcn_wave), 16, 0, 0);
}

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding)
#define NCVIS_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))

// Workgroup barrier without a memory fence: the stage pipelines order their LDS transfers with
// explicit vector-memory waits, and a fenced barrier would also wait for the transfers of the
// stages still in flight. The asm "memory" clobber keeps the compiler from moving LDS accesses
// across it.
#define NCVIS_BARRIER_NOFENCE() asm volatile("s_barrier" ::: "memory")

template <int P, int QB, int KL, int DK>
__global__ __launch_bounds__(256) void k_knn_tiles_dma(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                       int64_t N, int64_t Npad, float beta,
                                                       int32_t *__restrict__ cand) {
    using FT = typename Frag<P>::T;
    constexpr int ESZ = Frag<P>::ESZ;
    constexpr int DP = 32 * DK;
    constexpr int KS = (P == PREC_FP32) ? DP / 2 : DP / 16;
    constexpr int ROWB = DP * ESZ;
    constexpr int CB = (ROWB >= 512) ? 32 : 64;
    constexpr int CHUNKS = ROWB / 16;
    constexpr int NV = CB * ROWB / 16 / 256;
    constexpr int SB = CB * ROWB;
    static_assert(NV >= 1 && NV * 4096 == SB, "stage must split into whole wave transfers");
    static_assert((CHUNKS & (CHUNKS - 1)) == 0, "swizzle needs a power-of-two chunk count");

    __shared__ __align__(16) unsigned char lds0[SB];
    __shared__ __align__(16) unsigned char lds1[SB];
    __shared__ __align__(16) unsigned char lds2[SB];
    __shared__ __align__(16) float ldc0[4][CB];
    __shared__ __align__(16) float ldc1[4][CB];
    __shared__ __align__(16) float ldc2[4][CB];

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int h = lane >> 5;
    const int col = lane & 31;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * (32 * QB);

    FT qf[QB][KS];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        const unsigned char *row = Xh + (q0 + qb * 32 + col) * (int64_t)ROWB;
        if constexpr (P == PREC_FP32) {
#pragma unroll
            for (int s = 0; s < KS; s += 4) {
                float4 v = *(const float4 *)(row + (h * (DP / 2) + s) * 4);
                qf[qb][s] = v.x;
                qf[qb][s + 1] = v.y;
                qf[qb][s + 2] = v.z;
                qf[qb][s + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int s = 0; s < KS; ++s) qf[qb][s] = *(const uint4 *)(row + s * 32 + h * 16);
        }
    }
    float ld[QB][KL];
    int li[QB][KL];
    float thr[QB];
#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
#pragma unroll
        for (int p = 0; p < KL; ++p) {
            ld[qb][p] = INFINITY;
            li[qb][p] = -1;
        }
        thr[qb] = INFINITY;
    }

    const int64_t nstages = Npad / CB;
    const int64_t last = nstages - 1;
    const int swz = col & (CHUNKS - 1);  // rows rb*32 + col share col's swizzle (CHUNKS <= 32 divides 32)

#define NCVIS_DMA(STG, LB, LC) \
    dma_stage<NV, CB, ROWB, CHUNKS>(Xh, cn, (STG) < last ? (STG) : last, LB, LC[wave], wave, lane)

#define NCVIS_DMA_COMPUTE(LB, LC, ST)                                                                  \
    do {                                                                                               \
        _Pragma("unroll") for (int rb = 0; rb < CB / 32; ++rb) {                                       \
            const int64_t c0 = (ST) * CB + rb * 32;                                                    \
            const unsigned char *arow = LB + (rb * 32 + col) * ROWB;                                   \
            f32x16_t acc[QB];                                                                          \
            _Pragma("unroll") for (int qb = 0; qb < QB; ++qb) acc[qb] = f32x16_t{};                    \
            if constexpr (P == PREC_FP32) {                                                            \
                _Pragma("unroll") for (int s = 0; s < KS; s += 4) {                                    \
                    const int ch = (h * (DP / 2) + s) / 4;                                             \
                    float4 a4 = *(const float4 *)(arow + ((ch ^ swz) * 16));                          \
                    _Pragma("unroll") for (int qb = 0; qb < QB; ++qb) {                                \
                        acc[qb] = mfma_step<P>(a4.x, qf[qb][s], acc[qb]);                              \
                        acc[qb] = mfma_step<P>(a4.y, qf[qb][s + 1], acc[qb]);                          \
                        acc[qb] = mfma_step<P>(a4.z, qf[qb][s + 2], acc[qb]);                          \
                        acc[qb] = mfma_step<P>(a4.w, qf[qb][s + 3], acc[qb]);                          \
                    }                                                                                  \
                }                                                                                      \
            } else {                                                                                   \
                _Pragma("unroll") for (int s = 0; s < KS; ++s) {                                       \
                    uint4 a = *(const uint4 *)(arow + (((2 * s + h) ^ swz) * 16));                    \
                    _Pragma("unroll") for (int qb = 0; qb < QB; ++qb) acc[qb] =                        \
                        mfma_step<P>(a, qf[qb][s], acc[qb]);                                           \
                }                                                                                      \
            }                                                                                          \
            float cnv[16];                                                                             \
            _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                            \
                float4 v = *(const float4 *)(&LC[wave][rb * 32 + 8 * g + 4 * h]);                      \
                cnv[4 * g] = v.x;                                                                      \
                cnv[4 * g + 1] = v.y;                                                                  \
                cnv[4 * g + 2] = v.z;                                                                  \
                cnv[4 * g + 3] = v.w;                                                                  \
            }                                                                                          \
            _Pragma("unroll") for (int qb = 0; qb < QB; ++qb) {                                        \
                float m = INFINITY;                                                                    \
                _Pragma("unroll") for (int i = 0; i < 16; ++i) m = fminf(m, fmaf(beta, acc[qb][i], cnv[i])); \
                if (__any(m < thr[qb])) {                                                              \
                    _Pragma("unroll") for (int i = 0; i < 16; ++i) {                                   \
                        float sc = fmaf(beta, acc[qb][i], cnv[i]);                                     \
                        bool pass = sc < thr[qb];                                                      \
                        if (__any(pass)) {                                                             \
                            if (pass) {                                                                \
                                int ci = (int)(c0 + 8 * (i >> 2) + 4 * h + (i & 3));                   \
                                list_insert<KL>(ld[qb], li[qb], sc, ci);                               \
                                thr[qb] = ld[qb][KL - 1];                                              \
                            }                                                                          \
                        }                                                                              \
                    }                                                                                  \
                }                                                                                      \
            }                                                                                          \
        }                                                                                              \
    } while (0)

    // one step: wait for stage ST (only the next stage's transfers may still be in flight),
    // barrier, start stage ST+2 into the buffer stage ST-1 used, then compute stage ST
#define NCVIS_DMA_STEP(ST, LB, LC, NLB, NLC)                                                           \
    do {                                                                                               \
        NCVIS_WAIT_VMCNT(NV + 1);                                                                      \
        NCVIS_BARRIER_NOFENCE();                                                                       \
        NCVIS_DMA((ST) + 2, NLB, NLC);                                                                 \
        NCVIS_DMA_COMPUTE(LB, LC, ST);                                                                 \
    } while (0)

    NCVIS_DMA(0, lds0, ldc0);
    NCVIS_DMA(1, lds1, ldc1);
    for (int64_t st = 0; st < nstages; st += 3) {
        NCVIS_DMA_STEP(st, lds0, ldc0, lds2, ldc2);
        if (st + 1 >= nstages) break;
        NCVIS_DMA_STEP(st + 1, lds1, ldc1, lds0, ldc0);
        if (st + 2 >= nstages) break;
        NCVIS_DMA_STEP(st + 2, lds2, ldc2, lds1, ldc1);
    }
    NCVIS_WAIT_VMCNT(0);
#undef NCVIS_DMA_STEP
#undef NCVIS_DMA_COMPUTE
#undef NCVIS_DMA

#pragma unroll
    for (int qb = 0; qb < QB; ++qb) {
        int64_t q = q0 + qb * 32 + col;
        if (q < N) {
            int32_t *o = cand + q * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = li[qb][p];
        }
    }
}

// ------------------------------------------------------------------------ exact distances
__device__ inline float exact_dist(const float *__restrict__ x, const float *__restrict__ y, int D, int is_l2) {
    float s = 0.f;
    if (is_l2) {
        for (int j = 0; j < D; ++j) {
            float t = __fsub_rn(x[j], y[j]);
            s = __fadd_rn(s, __fmul_rn(t, t));
        }
        return s;
    }
    for (int j = 0; j < D; ++j) s = __fadd_rn(s, __fmul_rn(x[j], y[j]));
    return __fsub_rn(1.0f, s);
}

__device__ inline bool key_less(float da, int ia, int la, float db, int ib, int lb) {
    return da < db || (da == db && (ia < ib || (ia == ib && la < lb)));
}

// Exact distance with row loads four floats at a time; the sum stays sequential and uncontracted
// (same value as exact_dist). Rows must be 16-byte aligned (D % 4 == 0 and an aligned base).
__device__ inline float exact_dist4(const float *__restrict__ x, const float *__restrict__ y, int D, int is_l2) {
    const float4 *x4 = (const float4 *)x;
    const float4 *y4 = (const float4 *)y;
    float s = 0.f;
    if (is_l2) {
        for (int j = 0; j < D / 4; ++j) {
            const float4 a = x4[j], b = y4[j];
            float t = __fsub_rn(a.x, b.x);
            s = __fadd_rn(s, __fmul_rn(t, t));
            t = __fsub_rn(a.y, b.y);
            s = __fadd_rn(s, __fmul_rn(t, t));
            t = __fsub_rn(a.z, b.z);
            s = __fadd_rn(s, __fmul_rn(t, t));
            t = __fsub_rn(a.w, b.w);
            s = __fadd_rn(s, __fmul_rn(t, t));
        }
        return s;
    }
    for (int j = 0; j < D / 4; ++j) {
        const float4 a = x4[j], b = y4[j];
        s = __fadd_rn(s, __fmul_rn(a.x, b.x));
        s = __fadd_rn(s, __fmul_rn(a.y, b.y));
        s = __fadd_rn(s, __fmul_rn(a.z, b.z));
        s = __fadd_rn(s, __fmul_rn(a.w, b.w));
    }
    return __fsub_rn(1.0f, s);
}

// C = 2*KL candidates per query (32 or 64); a wave reranks 64/C queries, lane c of a query's
// segment handles candidate c: exact FP32 distance, rank by (distance, label, slot) with
// shuffles inside the segment, drop rank 0, write ranks k..1 (farthest first).
template <int C>
__global__ __launch_bounds__(256) void k_knn_rerank(const float *__restrict__ Xp, int64_t N, int D, int is_l2,
                                                    int vec4, const int32_t *__restrict__ cand, int k,
                                                    int32_t *__restrict__ inds, float *__restrict__ dists,
                                                    int *__restrict__ err) {
    constexpr int G = 64 / C;
    const int lane = threadIdx.x & 63;
    const int c = lane % C;
    const int64_t q = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * G + lane / C;
    const bool valid = q < N;
    int ci = valid ? cand[q * C + c] : -1;
    float dv = INFINITY;
    if (ci >= 0) {
        dv = vec4 ? exact_dist4(Xp + q * D, Xp + (int64_t)ci * D, D, is_l2)
                  : exact_dist(Xp + q * D, Xp + (int64_t)ci * D, D, is_l2);
        if (dv != dv) dv = INFINITY;  // NaN input: never ranks ahead of a real candidate
    } else {
        ci = INT_MAX;
    }
    int rank = 0;
#pragma unroll 8
    for (int j = 0; j < C; ++j) {
        const float dj = __shfl(dv, j, C);
        const int ij = __shfl(ci, j, C);
        rank += key_less(dj, ij, j, dv, ci, c) ? 1 : 0;
    }
    if (valid && rank >= 1 && rank <= k) {
        if (ci == INT_MAX) atomicExch(err, 1);
        inds[q * k + (k - rank)] = (ci == INT_MAX) ? 0 : ci;
        dists[q * k + (k - rank)] = dv;
    }
}

// 2*KL = 128 candidates per query (k + 1 up to 64): a wave per query, lanes l and l + 64 of the
// candidate list on each lane; same ranking and output as k_knn_rerank.
__global__ __launch_bounds__(256) void k_knn_rerank128(const float *__restrict__ Xp, int64_t N, int D, int is_l2,
                                                       int vec4, const int32_t *__restrict__ cand, int k,
                                                       int32_t *__restrict__ inds, float *__restrict__ dists,
                                                       int *__restrict__ err) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool valid = q < N;
    int ci[2];
    float dv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        ci[u] = valid ? cand[q * 128 + lane + 64 * u] : -1;
        dv[u] = INFINITY;
        if (ci[u] >= 0) {
            dv[u] = vec4 ? exact_dist4(Xp + q * D, Xp + (int64_t)ci[u] * D, D, is_l2)
                         : exact_dist(Xp + q * D, Xp + (int64_t)ci[u] * D, D, is_l2);
            if (dv[u] != dv[u]) dv[u] = INFINITY;
        } else {
            ci[u] = INT_MAX;
        }
    }
    int rank[2] = {0, 0};
    for (int j = 0; j < 64; ++j) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const float dj = __shfl(dv[v], j);
            const int ij = __shfl(ci[v], j);
#pragma unroll
            for (int u = 0; u < 2; ++u) rank[u] += key_less(dj, ij, j + 64 * v, dv[u], ci[u], lane + 64 * u) ? 1 : 0;
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        if (valid && rank[u] >= 1 && rank[u] <= k) {
            if (ci[u] == INT_MAX) atomicExch(err, 1);
            inds[q * k + (k - rank[u])] = (ci[u] == INT_MAX) ? 0 : ci[u];
            dists[q * k + (k - rank[u])] = dv[u];
        }
    }
}

// ------------------------------------------------------------------------ generic kNN
// Max-heap of the k+1 best (distance, label) pairs of one query ([k+1] in global scratch);
// returns the new entry count.
__device__ __noinline__ int heap_offer(float *hd, int *hi, int cnt, int K1, float dv, int ic) {
    if (cnt < K1) {
        int pos = cnt++;
        while (pos > 0) {
            int par = (pos - 1) >> 1;
            if (key_less(hd[par], hi[par], 0, dv, ic, 0)) {
                hd[pos] = hd[par];
                hi[pos] = hi[par];
                pos = par;
            } else {
                break;
            }
        }
        hd[pos] = dv;
        hi[pos] = ic;
    } else if (key_less(dv, ic, 0, hd[0], hi[0], 0)) {
        int pos = 0;
        while (true) {
            int l = 2 * pos + 1, r = l + 1, big = pos;
            float bd = dv;
            int bi = ic;
            if (l < K1 && key_less(bd, bi, 0, hd[l], hi[l], 0)) {
                big = l;
                bd = hd[l];
                bi = hi[l];
            }
            if (r < K1 && key_less(bd, bi, 0, hd[r], hi[r], 0)) big = r;
            if (big == pos) break;
            hd[pos] = hd[big];
            hi[pos] = hi[big];
            pos = big;
        }
        hd[pos] = dv;
        hi[pos] = ic;
    }
    return cnt;
}

// Exact FP32 scan for the shapes the MFMA tiles do not take (wide FP32 rows, rows over 512,
// k + 1 > 64): 256 queries per block, candidates in tiles of 64 and dimensions in chunks of 32
// staged in LDS (candidate values read as broadcasts), 64 running sums per thread. Each
// (query, candidate) sum runs over the dimensions in order with the same uncontracted operations
// as exact_dist, so every distance is exact_dist's. The k+1 best (distance, label) pairs of a
// query go through a max-heap, then the smallest is dropped and the rest written farthest first.
__global__ __launch_bounds__(256) void k_knn_generic(const float *__restrict__ Xp, int64_t N, int D, int is_l2,
                                                     int k, float *__restrict__ hd_, int *__restrict__ hi_,
                                                     int32_t *__restrict__ inds, float *__restrict__ dists) {
    constexpr int TC = 64, DC = 32;
    __shared__ float lq[256][DC + 1];
    __shared__ __align__(16) float lc[DC][TC];
    const int tid = threadIdx.x;
    const int64_t qb = (int64_t)blockIdx.x * 256;
    const int64_t q = qb + tid;
    const bool valid = q < N;
    const int K1 = k + 1;
    float *hd = hd_ + (valid ? q : 0) * K1;
    int *hi = hi_ + (valid ? q : 0) * K1;
    int cnt = 0;
    for (int64_t c0 = 0; c0 < N; c0 += TC) {
        float acc[TC];
#pragma unroll
        for (int t = 0; t < TC; ++t) acc[t] = 0.f;
        for (int d0 = 0; d0 < D; d0 += DC) {
            const int dn = (D - d0 < DC) ? D - d0 : DC;
            __syncthreads();
            for (int i = tid; i < 256 * DC; i += 256) {
                const int r = i / DC, c = i - r * DC;
                const int64_t row = qb + r;
                lq[r][c] = (row < N && c < dn) ? Xp[row * D + d0 + c] : 0.f;
            }
            for (int i = tid; i < TC * DC; i += 256) {
                const int t = i / DC, c = i - t * DC;
                const int64_t row = c0 + t;
                lc[c][t] = (row < N && c < dn) ? Xp[row * D + d0 + c] : 0.f;
            }
            __syncthreads();
            for (int d = 0; d < dn; ++d) {
                const float xq = lq[tid][d];
#pragma unroll
                for (int t4 = 0; t4 < TC; t4 += 4) {
                    const float4 y = *(const float4 *)&lc[d][t4];
                    const float yv[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (is_l2) {
                            const float e = __fsub_rn(xq, yv[u]);
                            acc[t4 + u] = __fadd_rn(acc[t4 + u], __fmul_rn(e, e));
                        } else {
                            acc[t4 + u] = __fadd_rn(acc[t4 + u], __fmul_rn(xq, yv[u]));
                        }
                    }
                }
            }
        }
        if (valid) {
            const int tn = (N - c0 < TC) ? (int)(N - c0) : TC;
#pragma unroll
            for (int t = 0; t < TC; ++t) {
                if (t < tn) {
                    float dv = is_l2 ? acc[t] : __fsub_rn(1.0f, acc[t]);
                    if (dv != dv) dv = INFINITY;
                    cnt = heap_offer(hd, hi, cnt, K1, dv, (int)(c0 + t));
                }
            }
        }
    }
    if (!valid) return;
    // heap-sort ascending in place, then drop the smallest and write farthest first
    for (int end = cnt - 1; end > 0; --end) {
        float td = hd[0];
        int ti = hi[0];
        float vd = hd[end];
        int vi = hi[end];
        hd[end] = td;
        hi[end] = ti;
        int pos = 0;
        while (true) {
            int l = 2 * pos + 1, r = l + 1, big = pos;
            float bd = vd;
            int bi = vi;
            if (l < end && key_less(bd, bi, 0, hd[l], hi[l], 0)) {
                big = l;
                bd = hd[l];
                bi = hi[l];
            }
            if (r < end && key_less(bd, bi, 0, hd[r], hi[r], 0)) big = r;
            if (big == pos) break;
            hd[pos] = hd[big];
            hi[pos] = hi[big];
            pos = big;
        }
        hd[pos] = vd;
        hi[pos] = vi;
    }
    for (int j = 0; j < k; ++j) {
        inds[q * k + j] = hi[K1 - 1 - j];
        dists[q * k + j] = hd[K1 - 1 - j];
    }
}

// ------------------------------------------------------------------------ host launcher
template <int P, int QB, int KL>
static void launch_tiles_dk(int DK, dim3 grid, const void *Xh, const float *cn, int64_t N, int64_t Npad,
                            float beta, int32_t *cand, hipStream_t st) {
    const Tuning &tu = tuning();
    const int select = tu.knn_select >= 0 ? tu.knn_select : 0;
    const int dma = tu.knn_dma >= 0 ? tu.knn_dma : (P == PREC_FP32 ? 1 : 0);
    if (select == 1) {
        DevBuf<uint2> sbuf((size_t)Npad * 2 * 4 * KL);
        switch (DK) {
            case 1: k_knn_tiles_buf<P, QB, KL, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, sbuf.p, cand); break;
            case 2: k_knn_tiles_buf<P, QB, KL, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, sbuf.p, cand); break;
            case 4: k_knn_tiles_buf<P, QB, KL, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, sbuf.p, cand); break;
            case 8: k_knn_tiles_buf<P, QB, KL, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, sbuf.p, cand); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
        NCVIS_LAUNCH_CHECK();
        NCVIS_HIP_CHECK(hipStreamSynchronize(st));  // sbuf is released on return
        return;
    }
    if (dma) {
        switch (DK) {
            case 1: k_knn_tiles_dma<P, QB, KL, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 2: k_knn_tiles_dma<P, QB, KL, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 4: k_knn_tiles_dma<P, QB, KL, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 8: k_knn_tiles_dma<P, QB, KL, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
        return;
    }
    if constexpr (P != PREC_FP32 && QB == 2 && KL == 16) {
        if (tu.knn_skew != 0) {
            if (tu.knn_stage_mult == 2) {
                switch (DK) {
                    case 1: k_knn_tiles_sk<P, KL, 1, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 2: k_knn_tiles_sk<P, KL, 2, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 4: k_knn_tiles_sk<P, KL, 4, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 8: k_knn_tiles_sk<P, KL, 8, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    default: throw HipError("[ncvis_amd] unsupported padded dimension");
                }
            } else {
                switch (DK) {
                    case 1: k_knn_tiles_sk<P, KL, 1, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 2: k_knn_tiles_sk<P, KL, 2, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 4: k_knn_tiles_sk<P, KL, 4, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    case 8: k_knn_tiles_sk<P, KL, 8, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                    default: throw HipError("[ncvis_amd] unsupported padded dimension");
                }
            }
            return;
        }
    }
    if (tu.knn_stage_mult == 2) {
        switch (DK) {
            case 1: k_knn_tiles<P, QB, KL, 1, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 2: k_knn_tiles<P, QB, KL, 2, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 4: k_knn_tiles<P, QB, KL, 4, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 8: k_knn_tiles<P, QB, KL, 8, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
        return;
    }
    switch (DK) {
        case 1: k_knn_tiles<P, QB, KL, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        case 2: k_knn_tiles<P, QB, KL, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        case 4: k_knn_tiles<P, QB, KL, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        case 8: k_knn_tiles<P, QB, KL, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        default: throw HipError("[ncvis_amd] unsupported padded dimension");
    }
}

template <int P>
static constexpr int qb_for() {
    return P == PREC_FP32 ? 1 : 2;
}

template <int P>
static void launch_tiles(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                         int32_t *cand, hipStream_t st) {
    constexpr int QB = qb_for<P>();
    dim3 grid((unsigned)(Npad / (4 * 32 * QB)));
    if (KL == 16)
        launch_tiles_dk<P, QB, 16>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
    else
        launch_tiles_dk<P, QB, 32>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
}

// k + 1 in 17..64 (lists of 32 or 64 per lane half): one query set per wave, since two sets'
// lists do not fit in VGPRs next to the query fragments (16-bit: register-staged tiles; FP32:
// the LDS-DMA tiles, which are one set per wave anyway).
template <int P, int KL>
static void launch_tiles_q1(int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                            int32_t *cand, hipStream_t st) {
    dim3 grid((unsigned)(Npad / (4 * 32)));
    if constexpr (P == PREC_FP32) {
        switch (DK) {
            case 1: k_knn_tiles_dma<P, 1, KL, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 2: k_knn_tiles_dma<P, 1, KL, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 4: k_knn_tiles_dma<P, 1, KL, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 8: k_knn_tiles_dma<P, 1, KL, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
    } else {
        switch (DK) {
            case 1: k_knn_tiles<P, 1, KL, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 2: k_knn_tiles<P, 1, KL, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 4: k_knn_tiles<P, 1, KL, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            case 8: k_knn_tiles<P, 1, KL, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
    }
}

// Rows of 257..512 elements (16-bit GEMM): one query set per wave, so the query fragments
// (Dpad / 16 per lane) still fit in VGPRs; 32-row stages.
template <int P, int KL>
static void launch_tiles_wide_kl(int DK, dim3 grid, const void *Xh, const float *cn, int64_t N, int64_t Npad,
                                 float beta, int32_t *cand, hipStream_t st) {
    switch (DK) {
        case 10: k_knn_tiles<P, 1, KL, 10><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        case 12: k_knn_tiles<P, 1, KL, 12><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        case 16: k_knn_tiles<P, 1, KL, 16><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
        default: throw HipError("[ncvis_amd] unsupported padded dimension");
    }
}

template <int P>
static void launch_tiles_wide(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                              int32_t *cand, hipStream_t st) {
    dim3 grid((unsigned)(Npad / (4 * 32)));
    if (KL == 16)
        launch_tiles_wide_kl<P, 16>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
    else
        launch_tiles_wide_kl<P, 32>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
}

void knn_run(const float *dX, int64_t N, int D, int k, int dist, int prec, int32_t *d_inds, float *d_dists,
             hipStream_t st, KnnStats *stats) {
    if (N < 2 || D < 1 || k < 1 || k > N - 1) throw HipError("[ncvis_amd] knn: invalid sizes");
    if (N > (int64_t)INT_MAX - 512) throw HipError("[ncvis_amd] knn: more than 2^31 points");
    const int is_l2 = (dist == DIST_L2);
    DevBuf<float> xp;
    const float *Xp = dX;
    if (dist == DIST_COSINE || dist == DIST_CORRELATION) {
        xp.alloc((size_t)N * D);
        k_preprocess<<<cdiv(N, 256), 256, 0, st>>>(dX, N, D, dist, xp.p);
        NCVIS_LAUNCH_CHECK();
        Xp = xp.p;
    }
    DevBuf<int> err(1);
    NCVIS_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipEvent_t e0, e1, e2;
    NCVIS_HIP_CHECK(hipEventCreate(&e0));
    NCVIS_HIP_CHECK(hipEventCreate(&e1));
    NCVIS_HIP_CHECK(hipEventCreate(&e2));
    // 16-bit GEMMs tile rows up to 512 elements, FP32 up to 256; wider rows or k + 1 > 32 take
    // the generic exact scan
    const bool wide = (prec != PREC_FP32) && (D > 256) && (D <= 512);
    const int KL = (k + 1 <= 16) ? 16 : (k + 1 <= 32) ? 32 : 64;
    const bool tiled = (D <= 256 && k + 1 <= 64) || (wide && k + 1 <= 32);
    if (tiled) {
        const int DK = D <= 32 ? 1 : D <= 64 ? 2 : D <= 128 ? 4 : D <= 256 ? 8 : D <= 320 ? 10 : D <= 384 ? 12 : 16;
        const int Dpad = 32 * DK;
        const int QB = (wide || KL >= 32) ? 1 : (prec == PREC_FP32) ? qb_for<PREC_FP32>() : qb_for<PREC_BF16>();
        const int64_t qtile = 4 * 32 * QB;
        const int64_t Npad = ((N + qtile - 1) / qtile) * qtile;
        const int esz = (prec == PREC_FP32) ? 4 : 2;
        DevBuf<unsigned char> xh((size_t)Npad * Dpad * esz);
        DevBuf<float> cn((size_t)Npad);
        switch (prec) {
            case PREC_FP32:
                k_pack<PREC_FP32><<<cdiv(Npad, 256), 256, 0, st>>>(Xp, N, Npad, D, Dpad, is_l2, xh.p, cn.p);
                break;
            case PREC_BF16:
                k_pack<PREC_BF16><<<cdiv(Npad, 256), 256, 0, st>>>(Xp, N, Npad, D, Dpad, is_l2, xh.p, cn.p);
                break;
            case PREC_FP16:
                k_pack<PREC_FP16><<<cdiv(Npad, 256), 256, 0, st>>>(Xp, N, Npad, D, Dpad, is_l2, xh.p, cn.p);
                break;
            default: throw HipError("[ncvis_amd] knn: unknown precision");
        }
        NCVIS_LAUNCH_CHECK();
        DevBuf<int32_t> cand((size_t)N * 2 * KL);
        const float beta = is_l2 ? -2.0f : -1.0f;
        NCVIS_HIP_CHECK(hipEventRecord(e0, st));
        if (KL == 64 && !wide) {
            switch (prec) {
                case PREC_FP32: launch_tiles_q1<PREC_FP32, 64>(DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
                case PREC_BF16: launch_tiles_q1<PREC_BF16, 64>(DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
                default: launch_tiles_q1<PREC_FP16, 64>(DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
            }
        } else if (KL == 32 && !wide && prec != PREC_FP32) {
            if (prec == PREC_BF16)
                launch_tiles_q1<PREC_BF16, 32>(DK, N, Npad, xh.p, cn.p, beta, cand.p, st);
            else
                launch_tiles_q1<PREC_FP16, 32>(DK, N, Npad, xh.p, cn.p, beta, cand.p, st);
        } else if (wide) {
            if (prec == PREC_BF16)
                launch_tiles_wide<PREC_BF16>(KL, DK, N, Npad, xh.p, cn.p, beta, cand.p, st);
            else
                launch_tiles_wide<PREC_FP16>(KL, DK, N, Npad, xh.p, cn.p, beta, cand.p, st);
        } else {
            switch (prec) {
                case PREC_FP32: launch_tiles<PREC_FP32>(KL, DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
                case PREC_BF16: launch_tiles<PREC_BF16>(KL, DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
                default: launch_tiles<PREC_FP16>(KL, DK, N, Npad, xh.p, cn.p, beta, cand.p, st); break;
            }
        }
        NCVIS_LAUNCH_CHECK();
        NCVIS_HIP_CHECK(hipEventRecord(e1, st));
        const int vec4 = (D % 4 == 0) && ((uintptr_t)Xp % 16 == 0);
        if (KL == 16)
            k_knn_rerank<32><<<cdiv(N, 8), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand.p, k, d_inds, d_dists, err.p);
        else if (KL == 32)
            k_knn_rerank<64><<<cdiv(N, 4), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand.p, k, d_inds, d_dists, err.p);
        else
            k_knn_rerank128<<<cdiv(N, 4), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand.p, k, d_inds, d_dists, err.p);
        NCVIS_LAUNCH_CHECK();
        NCVIS_HIP_CHECK(hipEventRecord(e2, st));
    } else {
        DevBuf<float> hd((size_t)N * (k + 1));
        DevBuf<int> hi((size_t)N * (k + 1));
        NCVIS_HIP_CHECK(hipEventRecord(e0, st));
        k_knn_generic<<<cdiv(N, 256), 256, 0, st>>>(Xp, N, D, is_l2, k, hd.p, hi.p, d_inds, d_dists);
        NCVIS_LAUNCH_CHECK();
        NCVIS_HIP_CHECK(hipEventRecord(e1, st));
        NCVIS_HIP_CHECK(hipEventRecord(e2, st));
    }
    int herr = 0;
    NCVIS_HIP_CHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (stats) {
        float t1 = 0, t2 = 0;
        NCVIS_HIP_CHECK(hipEventElapsedTime(&t1, e0, e1));
        NCVIS_HIP_CHECK(hipEventElapsedTime(&t2, e1, e2));
        stats->tile_ms = t1;
        stats->rerank_ms = t2;
        stats->used_mfma = tiled ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (herr) throw HipError("[ncvis_amd] knn: fewer than k+1 finite candidates (non-finite input?)");
}

}  // namespace ncvis_hip
