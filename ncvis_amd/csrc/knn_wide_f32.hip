// kNN candidate tiles for rows of 257..512 elements in the FP32 GEMM.
//
// The 32x32x2 FP32 tiles keep a query's Dpad / 2 operands per lane in VGPRs, which no longer fits
// beside the lists for rows this wide. Here each wave owns 16 queries and uses
// v_mfma_f32_16x16x4_f32: the four 16-lane groups of the wave split the reduction dimension, so a
// lane holds Dpad / 4 query operands. MFMA layout (16x16x4, one block): lane l supplies A[l % 16]
// [l / 16] and B[l / 16][l % 16] and receives D[4 (l / 16) + v][l % 16], v = 0..3. Candidates are
// the A rows and queries the B columns, so lane l scores candidates 4 (l / 16) .. +3 of each
// 16-row block for query l % 16, and keeps them in its own top-KL list: four quarter-lists per
// query, 4 KL candidates in all for the exact rerank.
//
// Rows are stored with element k = 4 t + g at position g (Dpad / 4) + t, so a lane's operands
// are contiguous. Candidate blocks stream through three LDS buffers filled by LDS DMA; inside each
// aligned group of 16 chunks of 16 bytes a row's chunks are XOR-swizzled with the row index, so
// the 16 rows a group of lanes reads at once fall in distinct banks (the row pitch is a multiple
// of the 64-bank width). Two accumulators take alternate k-steps so two MFMA chains are in flight.
#include "knn_tiles.h"

namespace ncvis_hip {

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// LDS slot of chunk c of row r (an involution: it also maps a slot back to its chunk)
__device__ __forceinline__ int wide_swizzle(int c, int r) { return (c & ~15) | ((c & 15) ^ (r & 15)); }

template <int CB, int CHUNKS, int NV>
__device__ __forceinline__ void dma_stage_wide(const float *__restrict__ Xh, const float *__restrict__ cn,
                                               int64_t stage, unsigned char *lbuf, float *lcn_wave, int wave,
                                               int lane) {
    const unsigned char *base = (const unsigned char *)Xh + stage * (int64_t)(CB * CHUNKS * 16);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const int p = (v * 4 + wave) * 64 + lane;  // 16-byte LDS slot this lane fills
        const int r = p / CHUNKS, s = p - r * CHUNKS;
        __builtin_amdgcn_global_load_lds(NCVIS_AS1(base + (r * CHUNKS + wide_swizzle(s, r)) * 16),
                                         NCVIS_AS3(lbuf + (v * 4 + wave) * 1024), 16, 0, 0);
    }
    if (lane < CB / 4)
        __builtin_amdgcn_global_load_lds(NCVIS_AS1(cn + stage * CB + 4 * lane), NCVIS_AS3(lcn_wave), 16, 0, 0);
}

template <int KL, int DK>
__global__ __launch_bounds__(256) void k_knn_tiles_f32w(const float *__restrict__ Xh, const float *__restrict__ cn,
                                                        int64_t N, int64_t Npad, float beta,
                                                        int32_t *__restrict__ cand) {
    constexpr int DP = 32 * DK;        // padded row length: 320, 384 or 512
    constexpr int KS = DP / 4;         // MFMA k-steps per row = query operands per lane
    constexpr int CB = 16;             // candidate rows per stage: one 16x16 block
    constexpr int CHUNKS = DP / 4;     // 16-byte chunks per row (DP * 4 / 16)
    constexpr int NV = CB * CHUNKS / 256;
    constexpr int SB = CB * CHUNKS * 16;
    static_assert(CHUNKS % 16 == 0, "swizzle groups of 16 chunks");
    static_assert(NV * 256 == CB * CHUNKS, "stage must split into whole wave transfers");
    static_assert(KS % 4 == 0, "operands are read four at a time");

    __shared__ __align__(16) unsigned char lds0[SB];
    __shared__ __align__(16) unsigned char lds1[SB];
    __shared__ __align__(16) unsigned char lds2[SB];
    __shared__ __align__(16) float ldc0[4][CB];
    __shared__ __align__(16) float ldc1[4][CB];
    __shared__ __align__(16) float ldc2[4][CB];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4;    // k group; also the candidate quarter of the block
    const int col = lane & 15;  // query column; also the candidate row this lane reads
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * 16;

    float qf[KS];
    {
        const float *qrow = Xh + (q0 + col) * DP + g * KS;
#pragma unroll
        for (int t = 0; t < KS; t += 4) {
            const float4 v = *(const float4 *)(qrow + t);
            qf[t] = v.x;
            qf[t + 1] = v.y;
            qf[t + 2] = v.z;
            qf[t + 3] = v.w;
        }
    }
    float ld[KL];
    int li[KL];
#pragma unroll
    for (int p = 0; p < KL; ++p) {
        ld[p] = INFINITY;
        li[p] = -1;
    }
    float thr = INFINITY;

    const int S = (int)gridDim.y, ysp = (int)blockIdx.y;  // candidate split (stages [sb, se))
    const int64_t total = Npad / CB;
    const int64_t sb = total * ysp / S, se = total * (ysp + 1) / S;
    const int64_t last = se - 1;

#define NCVIS_WDMA(STG, LB, LC) \
    dma_stage_wide<CB, CHUNKS, NV>(Xh, cn, (STG) < last ? (STG) : last, LB, LC[wave], wave, lane)

#define NCVIS_WCOMPUTE(LB, LC, ST)                                                                   \
    do {                                                                                             \
        const unsigned char *arow = (LB) + col * (CHUNKS * 16);                                      \
        f32x4_t acc0 = f32x4_t{}, acc1 = f32x4_t{};                                                  \
        _Pragma("unroll") for (int t = 0; t < KS; t += 4) {                                          \
            const int c = (g * KS + t) / 4;                                                          \
            const float4 a4 = *(const float4 *)(arow + wide_swizzle(c, col) * 16);                   \
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, qf[t], acc0, 0, 0, 0);                \
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, qf[t + 1], acc1, 0, 0, 0);            \
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, qf[t + 2], acc0, 0, 0, 0);            \
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, qf[t + 3], acc1, 0, 0, 0);            \
        }                                                                                            \
        const float4 cv = *(const float4 *)(&(LC)[wave][4 * g]);                                     \
        const float sc[4] = {fmaf(beta, acc0[0] + acc1[0], cv.x), fmaf(beta, acc0[1] + acc1[1], cv.y), \
                             fmaf(beta, acc0[2] + acc1[2], cv.z), fmaf(beta, acc0[3] + acc1[3], cv.w)}; \
        const float m = fminf(fminf(sc[0], sc[1]), fminf(sc[2], sc[3]));                             \
        if (__any(m < thr)) {                                                                        \
            _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                          \
                const bool pass = sc[v] < thr;                                                       \
                if (__any(pass)) {                                                                   \
                    if (pass) {                                                                      \
                        list_insert<KL>(ld, li, sc[v], (int)((ST) * CB + 4 * g + v));                \
                        thr = ld[KL - 1];                                                            \
                    }                                                                                \
                }                                                                                    \
            }                                                                                        \
        }                                                                                            \
    } while (0)

#define NCVIS_WSTEP(ST, LB, LC, NLB, NLC) \
    do {                                  \
        NCVIS_WAIT_VMCNT(NV + 1);         \
        NCVIS_BARRIER_NOFENCE();          \
        NCVIS_WDMA((ST) + 2, NLB, NLC);   \
        NCVIS_WCOMPUTE(LB, LC, ST);       \
    } while (0)

    NCVIS_WDMA(sb, lds0, ldc0);
    NCVIS_WDMA(sb + 1, lds1, ldc1);
    for (int64_t st = sb; st < se; st += 3) {
        NCVIS_WSTEP(st, lds0, ldc0, lds2, ldc2);
        if (st + 1 >= se) break;
        NCVIS_WSTEP(st + 1, lds1, ldc1, lds0, ldc0);
        if (st + 2 >= se) break;
        NCVIS_WSTEP(st + 2, lds2, ldc2, lds1, ldc1);
    }
    NCVIS_WAIT_VMCNT(0);
#undef NCVIS_WSTEP
#undef NCVIS_WCOMPUTE
#undef NCVIS_WDMA

    const int64_t q = q0 + col;
    if (q < N) {
        int32_t *o = cand + q * (4 * KL * S) + ysp * (4 * KL) + g * KL;
#pragma unroll
        for (int p = 0; p < KL; ++p) o[p] = li[p];
    }
}

void knn_tiles_wide_f32(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                        int32_t *cand, hipStream_t st, int split) {
    if (Npad % 64 != 0) throw HipError("[ncvis_amd] wide FP32 tiles need Npad % 64 == 0");
    const dim3 grid((unsigned)(Npad / 64), (unsigned)split);
    const float *X = (const float *)Xh;
#define NCVIS_W(KLV, DKV) k_knn_tiles_f32w<KLV, DKV><<<grid, 256, 0, st>>>(X, cn, N, Npad, beta, cand)
    if (KL == 16) {
        switch (DK) {
            case 10: NCVIS_W(16, 10); break;
            case 12: NCVIS_W(16, 12); break;
            case 16: NCVIS_W(16, 16); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
    } else if (KL == 32) {
        switch (DK) {
            case 10: NCVIS_W(32, 10); break;
            case 12: NCVIS_W(32, 12); break;
            case 16: NCVIS_W(32, 16); break;
            default: throw HipError("[ncvis_amd] unsupported padded dimension");
        }
    } else {
        throw HipError("[ncvis_amd] unsupported list length");
    }
#undef NCVIS_W
}

}  // namespace ncvis_hip
