// Device templates of the kNN candidate tiles (MFMA GEMM + per-lane top-KL lists) and their host
// launchers, shared by the translation units that instantiate them (knn_t16.hip, knn_q1_*.hip,
// knn_wide.hip); split so the heavy instantiations compile in parallel.
#pragma once

#include <cfloat>
#include <climits>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

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

    // candidate split: block row y of S sweeps stages [sb, se) and keeps lists of its own
    const int S = (int)gridDim.y, ysp = (int)blockIdx.y;
    const int64_t total = Npad / CB;
    const int64_t sb = total * ysp / S, se = total * (ysp + 1) / S;
    const bool cn_lane = threadIdx.x < CB / 4;  // these threads carry the stage's cn (float4 each)
    uint4 stage[NV];
    float4 stage_cn = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int v = 0; v < NV; ++v)
        stage[v] = *(const uint4 *)(Xh + sb * CB * (int64_t)ROWB + (int64_t)(v * 256 + threadIdx.x) * 16);
    if (cn_lane) stage_cn = *(const float4 *)(cn + sb * CB + 4 * threadIdx.x);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        int u = v * 256 + threadIdx.x;
        int r = u / CHUNKS, c = u - r * CHUNKS;
        *(uint4 *)(&lds[0][r * LROWB + c * 16]) = stage[v];
    }
    if (cn_lane) *(float4 *)(&ldc[0][4 * threadIdx.x]) = stage_cn;
    __syncthreads();

    for (int64_t st = sb; st < se; ++st) {
        const int buf = (int)((st - sb) & 1);
        {
            // next stage (the last iteration reloads its own stage; the copy is never read)
            const int64_t nx = (st + 1 < se) ? st + 1 : st;
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
                // minima of the lane's four groups of four scores: the insertion walks only the
                // groups in which some lane beats its (falling) threshold, in the same order
                float m[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    m[g] = INFINITY;
#pragma unroll
                    for (int u = 0; u < 4; ++u) m[g] = fminf(m[g], fmaf(beta, acc[qb][4 * g + u], cnv[4 * g + u]));
                }
                if (__any(fminf(fminf(m[0], m[1]), fminf(m[2], m[3])) < thr[qb])) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        if (!__any(m[g] < thr[qb])) continue;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            float sc = fmaf(beta, acc[qb][4 * g + u], cnv[4 * g + u]);
                            bool pass = sc < thr[qb];
                            if (__any(pass)) {
                                if (pass) {
                                    int ci = (int)(c0 + 8 * g + 4 * h + u);
                                    list_insert<KL>(ld[qb], li[qb], sc, ci);
                                    thr[qb] = ld[qb][KL - 1];
                                }
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
            int32_t *o = cand + q * (2 * KL * S) + ysp * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = li[qb][p];
        }
    }
}

// global / LDS address-space casts for __builtin_amdgcn_global_load_lds
#define NCVIS_AS1(p) ((const void __attribute__((address_space(1))) *)(p))
#define NCVIS_AS3(p) ((void __attribute__((address_space(3))) *)(p))

// One LDS-DMA stage: every lane moves 16-byte chunks of the stage's rows (swizzled so the
// fragment reads of a row land in distinct banks) and one wave-copy of the score offsets.
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

    const int S = (int)gridDim.y, ysp = (int)blockIdx.y;
    const int64_t total = Npad / CB;
    const int64_t sb = total * ysp / S, se = total * (ysp + 1) / S;
    const int64_t last = se - 1;
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
                float m[4];                                                                            \
                _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                        \
                    m[g] = INFINITY;                                                                   \
                    _Pragma("unroll") for (int u = 0; u < 4; ++u) m[g] =                               \
                        fminf(m[g], fmaf(beta, acc[qb][4 * g + u], cnv[4 * g + u]));                   \
                }                                                                                      \
                if (__any(fminf(fminf(m[0], m[1]), fminf(m[2], m[3])) < thr[qb])) {                    \
                    _Pragma("unroll") for (int g = 0; g < 4; ++g) {                                    \
                        if (!__any(m[g] < thr[qb])) continue;                                          \
                        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                \
                            float sc = fmaf(beta, acc[qb][4 * g + u], cnv[4 * g + u]);                 \
                            bool pass = sc < thr[qb];                                                  \
                            if (__any(pass)) {                                                         \
                                if (pass) {                                                            \
                                    int ci = (int)(c0 + 8 * g + 4 * h + u);                            \
                                    list_insert<KL>(ld[qb], li[qb], sc, ci);                           \
                                    thr[qb] = ld[qb][KL - 1];                                          \
                                }                                                                      \
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

    NCVIS_DMA(sb, lds0, ldc0);
    NCVIS_DMA(sb + 1, lds1, ldc1);
    for (int64_t st = sb; st < se; st += 3) {
        NCVIS_DMA_STEP(st, lds0, ldc0, lds2, ldc2);
        if (st + 1 >= se) break;
        NCVIS_DMA_STEP(st + 1, lds1, ldc1, lds0, ldc0);
        if (st + 2 >= se) break;
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
            int32_t *o = cand + q * (2 * KL * S) + ysp * (2 * KL) + h * KL;
#pragma unroll
            for (int p = 0; p < KL; ++p) o[p] = li[qb][p];
        }
    }
}

// ------------------------------------------------------------------------ host launcher
// k + 1 <= 16. 16-bit: two query sets per wave, LDS stages of 64 or 128 rows (knn_stage_mult);
// FP32: one set per wave, LDS-DMA stages by default (knn_dma = 0: register-staged).
template <int P>
static void launch_tiles16(int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                           int32_t *cand, hipStream_t st, const TileVariant &tv) {
    if constexpr (P == PREC_FP32) {
        dim3 grid((unsigned)(Npad / (4 * 32)), (unsigned)tv.split);
        if (tv.dma != 0) {
            switch (DK) {
                case 1: k_knn_tiles_dma<P, 1, 16, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 2: k_knn_tiles_dma<P, 1, 16, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 4: k_knn_tiles_dma<P, 1, 16, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 8: k_knn_tiles_dma<P, 1, 16, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                default: throw HipError("[ncvis_amd] unsupported padded dimension");
            }
        } else {
            switch (DK) {
                case 1: k_knn_tiles<P, 1, 16, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 2: k_knn_tiles<P, 1, 16, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 4: k_knn_tiles<P, 1, 16, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 8: k_knn_tiles<P, 1, 16, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                default: throw HipError("[ncvis_amd] unsupported padded dimension");
            }
        }
    } else {
        dim3 grid((unsigned)(Npad / (4 * 32 * 2)), (unsigned)tv.split);
        if (tv.stage_mult == 2) {
            switch (DK) {
                case 1: k_knn_tiles<P, 2, 16, 1, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 2: k_knn_tiles<P, 2, 16, 2, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 4: k_knn_tiles<P, 2, 16, 4, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 8: k_knn_tiles<P, 2, 16, 8, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                default: throw HipError("[ncvis_amd] unsupported padded dimension");
            }
        } else {
            switch (DK) {
                case 1: k_knn_tiles<P, 2, 16, 1><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 2: k_knn_tiles<P, 2, 16, 2><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 4: k_knn_tiles<P, 2, 16, 4><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                case 8: k_knn_tiles<P, 2, 16, 8><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand); break;
                default: throw HipError("[ncvis_amd] unsupported padded dimension");
            }
        }
    }
}

// k + 1 in 17..64 (lists of 32 or 64 per lane half): one query set per wave, since two sets'
// lists do not fit in VGPRs next to the query fragments. DMA = 1: LDS-DMA stages (FP32 only);
// DMA = 0: register-staged stages.
template <int P, int DMA>
static void launch_tiles_q1(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                            int32_t *cand, hipStream_t st, int split) {
    dim3 grid((unsigned)(Npad / (4 * 32)), (unsigned)split);
#define NCVIS_Q1(KLV, DKV)                                                                            \
    do {                                                                                             \
        if constexpr (DMA)                                                                           \
            k_knn_tiles_dma<P, 1, KLV, DKV><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand);     \
        else                                                                                         \
            k_knn_tiles<P, 1, KLV, DKV><<<grid, 256, 0, st>>>(Xh, cn, N, Npad, beta, cand);         \
    } while (0)
    if (KL != 32 && KL != 64) throw HipError("[ncvis_amd] unsupported list length");
    switch (DK) {
        case 1: if (KL == 32) NCVIS_Q1(32, 1); else NCVIS_Q1(64, 1); break;
        case 2: if (KL == 32) NCVIS_Q1(32, 2); else NCVIS_Q1(64, 2); break;
        case 4: if (KL == 32) NCVIS_Q1(32, 4); else NCVIS_Q1(64, 4); break;
        case 8: if (KL == 32) NCVIS_Q1(32, 8); else NCVIS_Q1(64, 8); break;
        default: throw HipError("[ncvis_amd] unsupported padded dimension");
    }
#undef NCVIS_Q1
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
                              int32_t *cand, hipStream_t st, int split) {
    dim3 grid((unsigned)(Npad / (4 * 32)), (unsigned)split);
    if (KL == 16)
        launch_tiles_wide_kl<P, 16>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
    else
        launch_tiles_wide_kl<P, 32>(DK, grid, Xh, cn, N, Npad, beta, cand, st);
}

}  // namespace ncvis_hip
