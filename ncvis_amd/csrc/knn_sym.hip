// Symmetric threshold search: the kNN candidate stage for 16-bit GEMMs, k + 1 <= 16, rows of at
// most 128 elements, computing every pair of rows once.
//
// The score of candidate c for query q is cn[c] + beta <q, c> and the one of q for c is
// cn[q] + beta <q, c>: one inner product serves both rows. The list tiles (knn_tiles.h) cannot use
// that, since the lists of the candidate side live in other workgroups' registers. Here every row
// first gets an upper bound T[q] on its (k+1)-th smallest score, then one pass over the upper
// triangle of query blocks emits every pair whose score is within either row's bound into that
// row's bucket, and each bucket is reduced to its 32 smallest (score, label) pairs:
//
//   k_sym_bound    - the list tiles' sweep over every stride-th row only (score-only half-lists of
//                    KL >= (k+1)/2); T[q] is the larger KL-th smallest of the two halves plus a
//                    margin sigma, an upper bound on the (k+1)-th smallest over all rows (2 KL rows
//                    of the subset score at most T[q])
//   k_sym_main     - a workgroup owns query blocks I and nb-1-I (balanced triangle halves) and
//                    sweeps the candidate stages from each block's own rows to the end. Pairs in
//                    the block's own stages are emitted for the query side only (every ordered pair
//                    is met there); pairs with a later candidate block are emitted for both rows
//   k_sym_select   - per row: the 32 smallest (score, label) pairs of its bucket (wave bitonic sort
//                    and merge), written in the rerank's candidate format; rows whose bucket
//                    overflowed or holds fewer than k+1 entries go to the fallback list
//   k_sym_fallback - the list tiles over the rows of the fallback list (gathered queries, a
//                    grid-stride loop over the list whose length is read on the device)
//
// A bucket receives every candidate whose score is at most T[q] (the query side tests the
// computed score exactly; the candidate side filters with a slack and then tests the computed
// score exactly), and T[q] is at least the (k+1)-th smallest score, so its 32 smallest pairs
// contain the k+1 smallest (score, label) pairs over all candidates: the exact rerank then writes
// the table the list tiles write on exactly scored data. T depends only on the data, so the
// buckets' contents, and the candidates the rerank sees, do not depend on the emission order.
#include "knn_tiles.h"

namespace ncvis_hip {

// stage chunks as a native vector type: copies stay register moves (a uint4 struct copy is a
// memcpy that can keep the whole stage array in scratch memory)
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

template <int DK>
struct SymShape {
    static constexpr int DP = 32 * DK;                          // padded row length (elements)
    static constexpr int KS = DP / 16;                          // MFMA k-steps per row
    static constexpr int ROWB = DP * 2;                         // 16-bit row bytes
    // candidate rows per LDS stage: 128 for rows up to 64 elements, 64 above (staging registers
    // beside the query fragments and the MFMA results stay within 256 VGPRs)
    static constexpr int CB = (ROWB <= 128) ? 128 : (ROWB >= 512 ? 32 : 64);
    static constexpr int LROWB = ROWB + 16;                     // padded LDS row (bank spread)
    static constexpr int CHUNKS = ROWB / 16;                    // 16-byte chunks per row
    static constexpr int NV = CB * ROWB / 16 / 256;             // chunks per thread per stage
    static constexpr int SPB = 256 / CB;                        // stages per query block (256 rows)
    static_assert(NV >= 1 && NV * 256 * 16 == CB * ROWB, "stage must split evenly over 256 threads");
};

// Sorted ascending score list without labels; requires x < ld[KL-1].
template <int KL>
__device__ __forceinline__ void score_insert(float (&ld)[KL], float x) {
#pragma unroll
    for (int p = KL - 1; p >= 1; --p) ld[p] = __builtin_amdgcn_fmed3f(x, ld[p - 1], ld[p]);
    ld[0] = fminf(x, ld[0]);
}

// Query fragments of rows r[qb] (two query sets of 32 rows per wave).
template <int DK>
__device__ __forceinline__ void sym_load_queries(const unsigned char *Xh, const int64_t (&r)[2], int h,
                                                 uint4 (&qf)[2][SymShape<DK>::KS]) {
    using Sh = SymShape<DK>;
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        const unsigned char *row = Xh + r[qb] * (int64_t)Sh::ROWB;
#pragma unroll
        for (int s = 0; s < Sh::KS; ++s) qf[qb][s] = *(const uint4 *)(row + s * 32 + h * 16);
    }
}

// The 32x32 block of candidate rows at arow (this lane's row, its half already added) against
// both query sets. MFMA layout: lane l holds query l % 32 and candidates 8 (i / 4) + 4 (l / 32)
// + (i % 4), i = 0..15.
template <int P, int DK>
__device__ __forceinline__ void sym_mfma(const unsigned char *arow, const uint4 (&qf)[2][SymShape<DK>::KS],
                                         f32x16_t (&acc)[2]) {
    acc[0] = f32x16_t{};
    acc[1] = f32x16_t{};
#pragma unroll
    for (int s = 0; s < SymShape<DK>::KS; ++s) {
        const uint4 a = *(const uint4 *)(arow + s * 32);
        acc[0] = mfma_step<P>(a, qf[0][s], acc[0]);
        acc[1] = mfma_step<P>(a, qf[1][s], acc[1]);
    }
}

template <int DK>
__device__ __forceinline__ void sym_store_stage(unsigned char *lbuf, const u32x4_t (&stage)[SymShape<DK>::NV]) {
    using Sh = SymShape<DK>;
#pragma unroll
    for (int v = 0; v < Sh::NV; ++v) {
        const int u = v * 256 + threadIdx.x;
        const int r = u / Sh::CHUNKS, c = u - r * Sh::CHUNKS;
        *(u32x4_t *)(lbuf + r * Sh::LROWB + c * 16) = stage[v];
    }
}

// contiguous candidate rows [r0, r0 + CB)
template <int DK>
__device__ __forceinline__ void sym_load_stage(const unsigned char *Xh, int64_t r0, u32x4_t (&stage)[SymShape<DK>::NV]) {
    using Sh = SymShape<DK>;
    const unsigned char *b = Xh + r0 * (int64_t)Sh::ROWB;
#pragma unroll
    for (int v = 0; v < Sh::NV; ++v) stage[v] = *(const u32x4_t *)(b + (int64_t)(v * 256 + threadIdx.x) * 16);
}

// Gathered candidate stage st: virtual candidates j = st * CB + r; j < n1 is row rows1[j], then
// n2 rows (j - n1) * stride2; beyond that padding (offset +inf, any row's data).
template <int DK>
__device__ __forceinline__ void sym_load_gathered(const unsigned char *Xh, const float *cn, int64_t st,
                                                  const int *rows1, int n1, int n2, int64_t stride2, bool cn_lane,
                                                  u32x4_t (&stage)[SymShape<DK>::NV], float4 &scn) {
    using Sh = SymShape<DK>;
#pragma unroll
    for (int v = 0; v < Sh::NV; ++v) {
        const int u = v * 256 + threadIdx.x;
        const int r = u / Sh::CHUNKS, c = u - r * Sh::CHUNKS;
        const int64_t j = st * Sh::CB + r;
        const int64_t row = (j < n1) ? rows1[j] : (j < n1 + n2 ? (j - n1) * stride2 : 0);
        stage[v] = *(const u32x4_t *)(Xh + row * Sh::ROWB + c * 16);
    }
    if (cn_lane) {
        float e[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t j = st * Sh::CB + 4 * threadIdx.x + t;
            e[t] = (j < n1) ? cn[rows1[j]] : (j < n1 + n2 ? cn[(j - n1) * stride2] : INFINITY);
        }
        scn = make_float4(e[0], e[1], e[2], e[3]);
    }
}

// ------------------------------------------------------------------------ row bounds
// Every stride-th row (j * stride, j < Ns) as candidates of every query, each lane half keeping the
// KL smallest scores of the candidates it sees. T[q] = the larger of the two halves' KL-th smallest
// + sigma: 2 KL distinct rows score at most T[q], so T[q] bounds the (2 KL)-th smallest score over
// all rows (+inf when a half saw fewer than KL finite scores); padding queries get -inf.
template <int P, int DK, int KL>
__global__ __launch_bounds__(256, 2) void k_sym_bound(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                      int64_t N, float beta, int64_t stride, int64_t Ns, float sigma,
                                                      float *__restrict__ Tq) {
    using Sh = SymShape<DK>;
    constexpr int CB = Sh::CB;
    __shared__ __align__(16) unsigned char lds[2][CB * Sh::LROWB];
    __shared__ __align__(16) float ldc[2][CB];

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int col = lane & 31;
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
    const int64_t qr[2] = {q0 + col, q0 + 32 + col};
    uint4 qf[2][Sh::KS];
    sym_load_queries<DK>(Xh, qr, h, qf);
    float ld[2][KL];
    float thr[2] = {INFINITY, INFINITY};
#pragma unroll
    for (int qb = 0; qb < 2; ++qb)
#pragma unroll
        for (int p = 0; p < KL; ++p) ld[qb][p] = INFINITY;

    const int64_t total = (Ns + CB - 1) / CB;
    const bool cn_lane = threadIdx.x < CB / 4;
    u32x4_t stage[Sh::NV];
    float4 scn = make_float4(0.f, 0.f, 0.f, 0.f);
    sym_load_gathered<DK>(Xh, cn, 0, nullptr, 0, (int)Ns, stride, cn_lane, stage, scn);
    sym_store_stage<DK>(lds[0], stage);
    if (cn_lane) *(float4 *)(&ldc[0][4 * threadIdx.x]) = scn;
    __syncthreads();
    for (int64_t st = 0; st < total; ++st) {
        const int buf = (int)(st & 1);
        sym_load_gathered<DK>(Xh, cn, st + 1 < total ? st + 1 : st, nullptr, 0, (int)Ns, stride, cn_lane, stage, scn);
#pragma unroll
        for (int rb = 0; rb < CB / 32; ++rb) {
            f32x16_t acc[2];
            sym_mfma<P, DK>(&lds[buf][(rb * 32 + col) * Sh::LROWB + h * 16], qf, acc);
            float cnv[16];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 v = *(const float4 *)(&ldc[buf][rb * 32 + 8 * g + 4 * h]);
                cnv[4 * g] = v.x;
                cnv[4 * g + 1] = v.y;
                cnv[4 * g + 2] = v.z;
                cnv[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int qb = 0; qb < 2; ++qb) {
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
                            const float sc = fmaf(beta, acc[qb][4 * g + u], cnv[4 * g + u]);
                            const bool pass = sc < thr[qb];
                            if (__any(pass)) {
                                if (pass) {
                                    score_insert<KL>(ld[qb], sc);
                                    thr[qb] = ld[qb][KL - 1];
                                }
                            }
                        }
                    }
                }
            }
        }
        sym_store_stage<DK>(lds[buf ^ 1], stage);
        if (cn_lane) *(float4 *)(&ldc[buf ^ 1][4 * threadIdx.x]) = scn;
        __syncthreads();
    }
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        const float T = fmaxf(ld[qb][KL - 1], __shfl_xor(ld[qb][KL - 1], 32));
        if (h == 0) Tq[qr[qb]] = (qr[qb] < N) ? T + sigma : -INFINITY;
    }
}

// ------------------------------------------------------------------------ emission pass
__device__ __forceinline__ void sym_emit(int *__restrict__ cnt, uint2 *__restrict__ bucket, int B, int row, int idx,
                                         float s) {
    const int slot = atomicAdd(&cnt[row], 1);
    if (slot < B) bucket[(int64_t)row * B + slot] = make_uint2(__float_as_uint(s), (unsigned)idx);
}

// Emissions of a workgroup queue in LDS (an LDS atomic reserves the slot) and move to the row
// buckets in bulk once enough have gathered: the 256 threads then issue their device-scope
// atomics together, so no wave waits on a bucket counter inside the sweep. A full queue sends an
// emission straight to its bucket. The queue has two halves used by alternate stages: the half of
// stage t only changes again in stage t + 2, after the barrier that ends stage t + 1, so every
// wave reads the same count after the barrier that ends stage t and the flush decision is
// uniform.
constexpr int kSymHalf = 256;        // LDS queue entries per half
constexpr int kSymHalfFlush = 128;   // flush a half once this many have gathered

struct SymQueue {
    int row[2][kSymHalf];
    uint2 ent[2][kSymHalf];
    int n[2];
};

__device__ __forceinline__ void sym_queue(SymQueue &q, int par, int *__restrict__ cnt, uint2 *__restrict__ bucket,
                                          int B, int row, int idx, float s) {
    const int slot = atomicAdd(&q.n[par], 1);
    if (slot < kSymHalf) {
        q.row[par][slot] = row;
        q.ent[par][slot] = make_uint2(__float_as_uint(s), (unsigned)idx);
    } else {
        sym_emit(cnt, bucket, B, row, idx, s);
    }
}

// every thread of the workgroup; n = the half's count read after a barrier (uniform)
__device__ __forceinline__ void sym_flush(SymQueue &q, int par, int n, int *__restrict__ cnt,
                                          uint2 *__restrict__ bucket, int B) {
    if (n > kSymHalf) n = kSymHalf;
    int slot[kSymHalf / 256];
#pragma unroll
    for (int u = 0; u < kSymHalf / 256; ++u) {
        const int e = u * 256 + threadIdx.x;
        slot[u] = (e < n) ? atomicAdd(&cnt[q.row[par][e]], 1) : B;
    }
#pragma unroll
    for (int u = 0; u < kSymHalf / 256; ++u) {
        const int e = u * 256 + threadIdx.x;
        if (slot[u] < B) bucket[(int64_t)q.row[par][e] * B + slot[u]] = q.ent[par][e];
    }
    __syncthreads();
    if (threadIdx.x == 0) q.n[par] = 0;
    __syncthreads();
}

// Stages of the emission pass arrive by LDS DMA (no staging registers): rows unpadded, the 16-byte
// chunks of row r at positions XOR-swizzled with r so a fragment read's 32 rows fall in distinct
// banks; two buffers, the next stage's transfer overlapping the current stage's MFMA work.
template <int DK>
struct SymDma {
    static constexpr int ROWB = 64 * DK;                  // 16-bit row bytes
    static constexpr int CB = (ROWB >= 512) ? 64 : 128;   // candidate rows per stage
    static constexpr int CHUNKS = ROWB / 16;
    static constexpr int NV = CB * ROWB / 4096;           // 16-byte transfers per lane per stage
    static constexpr int SPB = 256 / CB;
    static_assert(NV >= 1 && NV * 4096 == CB * ROWB, "stage must split into whole wave transfers");
    static_assert(CHUNKS <= 32 && (CHUNKS & (CHUNKS - 1)) == 0, "swizzle by the lane's row");
};

template <int DK>
__device__ __forceinline__ void sym_dma_stage(const unsigned char *__restrict__ Xh, const float *__restrict__ cn,
                                              const float *__restrict__ Tq, int64_t st, unsigned char *lbuf,
                                              float *lc, float *lt, int wave, int lane) {
    using D = SymDma<DK>;
    const unsigned char *base = Xh + st * (int64_t)(D::CB * D::ROWB);
#pragma unroll
    for (int v = 0; v < D::NV; ++v) {
        const int p = (v * 4 + wave) * 64 + lane;  // 16-byte LDS slot this lane fills
        const int r = p / D::CHUNKS, sl = p % D::CHUNKS;
        const int c = sl ^ (r & (D::CHUNKS - 1));
        __builtin_amdgcn_global_load_lds(NCVIS_AS1(base + r * D::ROWB + c * 16), NCVIS_AS3(lbuf + (v * 4 + wave) * 1024),
                                         16, 0, 0);
    }
    if (lane < D::CB / 4) {
        if (wave == 0) __builtin_amdgcn_global_load_lds(NCVIS_AS1(cn + st * D::CB + 4 * lane), NCVIS_AS3(lc), 16, 0, 0);
        if (wave == 1) __builtin_amdgcn_global_load_lds(NCVIS_AS1(Tq + st * D::CB + 4 * lane), NCVIS_AS3(lt), 16, 0, 0);
    }
}

// Work items of the chunk-major order: candidate chunk c (stages [c * chunk, (c + 1) * chunk))
// against every query block I whose own stages start before the chunk ends, I < (c + 1) * m with
// m = chunk / SPB query blocks per chunk; items run chunk by chunk, so the workgroups resident on
// an XCD at one time sweep the same chunk (chunk * CB rows, which the XCD's L2 holds) instead of
// each streaming its own range of the whole candidate matrix from memory.
__host__ __device__ inline int64_t sym_chunk_prefix(int64_t c, int64_t m, int64_t nbI) {
    const int64_t full = nbI / m;            // chunks c with (c + 1) m <= nbI
    const int64_t k = c < full ? c : full;   // of the first c chunks, those below the cap
    return m * k * (k + 1) / 2 + (c - k) * nbI;
}

template <int P, int DK>
__global__ __launch_bounds__(256, 2) void k_sym_main(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                     const float *__restrict__ Tq, int64_t N, int64_t Npad,
                                                     float beta, float sigma, int B, int *__restrict__ cnt,
                                                     uint2 *__restrict__ bucket, int chunk) {
    using D = SymDma<DK>;
    constexpr int CB = D::CB, SPB = D::SPB, KS = DK * 2;
    __shared__ __align__(16) unsigned char lds[2][CB * D::ROWB];
    __shared__ __align__(16) float ldc[2][CB];
    __shared__ __align__(16) float ldt[2][CB];
    __shared__ SymQueue sq;

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int h = lane >> 5;
    const int col = lane & 31;
    const int swz = col & (D::CHUNKS - 1);  // rows rb * 32 + col share col's swizzle

    const int64_t nbI = Npad / 256, NS = Npad / CB;
    int64_t I1, I2, len1, vb, ve, s_lo = 0;
    if (chunk > 0) {
        // chunk-major item: chunk c, query block I1, stages [max(c chunk, I1 SPB), min((c+1) chunk, NS))
        const int64_t m = chunk / SPB, nC = (NS + chunk - 1) / chunk, item = blockIdx.x;
        int64_t lo = 0, hi = nC;  // largest c with prefix(c) <= item
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (sym_chunk_prefix(mid, m, nbI) <= item)
                lo = mid;
            else
                hi = mid;
        }
        I1 = item - sym_chunk_prefix(lo, m, nbI);
        I2 = I1;
        s_lo = lo * chunk > I1 * SPB ? lo * chunk : I1 * SPB;
        const int64_t s_hi = (lo + 1) * chunk < NS ? (lo + 1) * chunk : NS;
        len1 = s_hi - s_lo;  // virtual stage v = stage s_lo + v, all of block I1
        vb = 0;
        ve = len1;
    } else {
        // the two query blocks of this workgroup and its share of their concatenated sweeps
        I1 = blockIdx.x;
        I2 = nbI - 1 - (int64_t)blockIdx.x;
        len1 = NS - I1 * SPB;
        const int64_t len2 = (I2 > I1) ? NS - I2 * SPB : 0;
        const int64_t Lt = len1 + len2;
        vb = Lt * blockIdx.y / gridDim.y;
        ve = Lt * (blockIdx.y + 1) / gridDim.y;
    }
    if (vb >= ve) return;  // uniform over the workgroup
    // pair order: virtual stage v is block I1's stages from the last one down to its own rows, then
    // block I2's (every workgroup sweeps the candidate stages in the same descending order);
    // chunk order: the item's stages ascending
#define NCVIS_SYM_STAGE(V) (chunk > 0 ? s_lo + (V) : ((V) < len1 ? NS - 1 - (V) : NS - 1 - ((V)-len1)))

    uint4 qf[2][KS];
    int qi[2] = {0, 0};
    float tq[2], cnq[2], tcs[2];
    int64_t cur = -1;
    const int Ni = (int)N;

    sym_dma_stage<DK>(Xh, cn, Tq, NCVIS_SYM_STAGE(vb), lds[0], ldc[0], ldt[0], wave, lane);
    if (threadIdx.x < 2) sq.n[threadIdx.x] = 0;
    NCVIS_WAIT_VMCNT(0);
    __syncthreads();

    for (int64_t v = vb; v < ve; ++v) {
        const int buf = (int)((v - vb) & 1);
        const int64_t blk = (v < len1) ? I1 : I2;
        const int64_t st = NCVIS_SYM_STAGE(v);
        if (blk != cur) {
            cur = blk;
            const int q0 = (int)blk * 256 + wave * 64;
            qi[0] = q0 + col;
            qi[1] = q0 + 32 + col;
            const int64_t qr[2] = {qi[0], qi[1]};
            sym_load_queries<DK>(Xh, qr, h, qf);
#pragma unroll
            for (int qb = 0; qb < 2; ++qb) {
                tq[qb] = Tq[qi[qb]];
                cnq[qb] = cn[qi[qb]];
                tcs[qb] = (qi[qb] < Ni) ? sigma - cnq[qb] : -INFINITY;
            }
        }
        // the next stage into the other buffer (read last in the previous stage, before the barrier)
        if (v + 1 < ve) sym_dma_stage<DK>(Xh, cn, Tq, NCVIS_SYM_STAGE(v + 1), lds[buf ^ 1], ldc[buf ^ 1], ldt[buf ^ 1],
                                          wave, lane);
        const bool diag = st < (blk + 1) * SPB;  // the query block's own rows: query side only
        // one 32-row block at a time keeps the live registers within 256
#pragma unroll 1
        for (int rb = 0; rb < CB / 32; ++rb) {
            const int c0 = (int)(st * CB) + rb * 32;  // rows < 2^31
            const unsigned char *arow = &lds[buf][(rb * 32 + col) * D::ROWB];
            f32x16_t acc[2];
            acc[0] = f32x16_t{};
            acc[1] = f32x16_t{};
#pragma unroll
            for (int s2 = 0; s2 < KS; ++s2) {
                const uint4 a = *(const uint4 *)(arow + (((2 * s2 + h) ^ swz) * 16));
                acc[0] = mfma_step<P>(a, qf[0][s2], acc[0]);
                acc[1] = mfma_step<P>(a, qf[1][s2], acc[1]);
            }
            float cnv[16], tv[16];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 a = *(const float4 *)(&ldc[buf][rb * 32 + 8 * g + 4 * h]);
                cnv[4 * g] = a.x;
                cnv[4 * g + 1] = a.y;
                cnv[4 * g + 2] = a.z;
                cnv[4 * g + 3] = a.w;
            }
            if (!diag) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float4 a = *(const float4 *)(&ldt[buf][rb * 32 + 8 * g + 4 * h]);
                    tv[4 * g] = a.x;
                    tv[4 * g + 1] = a.y;
                    tv[4 * g + 2] = a.z;
                    tv[4 * g + 3] = a.w;
                }
            }
#pragma unroll
            for (int qb = 0; qb < 2; ++qb) {
                // query side: score of candidate r for q is cn[r] + beta <q, r>. Candidate side:
                // score of q for r is cn[q] + beta <q, r>; filtered with beta <q, r> - T[r] <=
                // sigma - cn[q], then tested exactly
                // the minima of the lane's four groups of four candidates (i = 4 g .. 4 g + 3): the
                // emission walks only the groups in which some lane of the wave passes
                float m1[4], m2[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    m1[g] = INFINITY;
                    m2[g] = INFINITY;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = 4 * g + u;
                        m1[g] = fminf(m1[g], fmaf(beta, acc[qb][i], cnv[i]));
                        if (!diag) m2[g] = fminf(m2[g], fmaf(beta, acc[qb][i], -tv[i]));
                    }
                }
                if (__any(fminf(fminf(m1[0], m1[1]), fminf(m1[2], m1[3])) <= tq[qb])) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        if (!__any(m1[g] <= tq[qb])) continue;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int i = 4 * g + u;
                            const int off = 8 * g + 4 * h + u;
                            const float s1 = fmaf(beta, acc[qb][i], cnv[i]);
                            if (s1 <= tq[qb] && c0 + off < Ni && qi[qb] < Ni)
                                sym_queue(sq, buf, cnt, bucket, B, qi[qb], c0 + off, s1);
                        }
                    }
                }
                if (!diag && __any(fminf(fminf(m2[0], m2[1]), fminf(m2[2], m2[3])) <= tcs[qb])) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        if (!__any(m2[g] <= tcs[qb])) continue;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int i = 4 * g + u;
                            const int off = 8 * g + 4 * h + u;
                            if (fmaf(beta, acc[qb][i], -tv[i]) <= tcs[qb]) {
                                const float s2 = fmaf(beta, acc[qb][i], cnq[qb]);
                                // padding queries (cn = +inf) would pass against an unbounded row
                                if (c0 + off < Ni && qi[qb] < Ni && s2 <= tv[i])
                                    sym_queue(sq, buf, cnt, bucket, B, c0 + off, qi[qb], s2);
                            }
                        }
                    }
                }
            }
        }
        NCVIS_WAIT_VMCNT(0);  // this lane's transfers of the next stage
        __syncthreads();
        const int nq = sq.n[buf];  // uniform: this half takes no entries before the next barrier
        if (nq >= kSymHalfFlush) sym_flush(sq, buf, nq, cnt, bucket, B);
    }
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int nq = sq.n[par];  // the sweep is over: both halves are final
        if (nq > 0) sym_flush(sq, par, nq, cnt, bucket, B);
    }
#undef NCVIS_SYM_STAGE
}

// ------------------------------------------------------------------------ bucket reduction
__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long x, int m) {
    const unsigned lo = __shfl_xor((unsigned)x, m), hi = __shfl_xor((unsigned)(x >> 32), m);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl64(unsigned long long x, int src) {
    const unsigned lo = __shfl((unsigned)x, src), hi = __shfl((unsigned)(x >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}

// (score, label) as one ascending 64-bit key: order-preserving float bits, then the label
__device__ __forceinline__ unsigned long long sym_key(uint2 e) {
    const unsigned u = (e.x & 0x80000000u) ? ~e.x : (e.x | 0x80000000u);
    return ((unsigned long long)u << 32) | e.y;
}

__device__ __forceinline__ unsigned long long kmin(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
__device__ __forceinline__ unsigned long long kmax(unsigned long long a, unsigned long long b) { return a < b ? b : a; }

// one wave per row: the 32 smallest keys of the bucket, or the row goes to the fallback list
__global__ __launch_bounds__(256) void k_sym_select(const int *__restrict__ cnt, const uint2 *__restrict__ bucket,
                                                    int B, int64_t N, int K1, int32_t *__restrict__ cand,
                                                    int *__restrict__ ovf, int *__restrict__ n_ovf) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= N) return;  // uniform over the wave
    const int c = cnt[q];
    if (c > B || c < K1) {
        if (lane == 0) ovf[atomicAdd(n_ovf, 1)] = (int)q;
        return;
    }
    const unsigned long long EMPTY = ~0ull;
    unsigned long long top = EMPTY;  // lanes 0..31: the smallest keys so far, ascending
    for (int base = 0; base < c; base += 64) {
        unsigned long long k = (base + lane < c) ? sym_key(bucket[q * B + base + lane]) : EMPTY;
        // bitonic sort of the chunk over the wave, ascending
        for (int size = 2; size <= 64; size <<= 1) {
            for (int s = size >> 1; s > 0; s >>= 1) {
                const unsigned long long o = shfl_xor64(k, s);
                const bool up = (lane & size) == 0 || size == 64;
                const bool lower = (lane & s) == 0;
                k = (lower == up) ? kmin(k, o) : kmax(k, o);
            }
        }
        // merge with the running list: ascending top in lanes 0..31 and the chunk's smallest 32
        // descending in lanes 32..63 form a bitonic sequence
        const unsigned long long rev = shfl64(k, 63 - lane);
        unsigned long long m = (lane < 32) ? top : rev;
        for (int s = 32; s > 0; s >>= 1) {
            const unsigned long long o = shfl_xor64(m, s);
            m = ((lane & s) == 0) ? kmin(m, o) : kmax(m, o);
        }
        top = (lane < 32) ? m : EMPTY;
    }
    if (lane < 32) cand[q * 32 + lane] = (top == EMPTY) ? -1 : (int32_t)(unsigned)(top & 0xffffffffu);
}

// ------------------------------------------------------------------------ fallback rows
// The list tiles (lists of 16 per lane half) for the rows of `list`, writing their 32 candidates.
template <int P, int DK>
__global__ __launch_bounds__(256, 2) void k_sym_fallback(const void *__restrict__ Xh_, const float *__restrict__ cn,
                                                      int64_t Npad, float beta, const int *__restrict__ list,
                                                      const int *__restrict__ n_list, int32_t *__restrict__ cand) {
    using Sh = SymShape<DK>;
    constexpr int CB = Sh::CB;
    __shared__ __align__(16) unsigned char lds[2][CB * Sh::LROWB];
    __shared__ __align__(16) float ldc[2][CB];

    const unsigned char *Xh = (const unsigned char *)Xh_;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = lane >> 5;
    const int col = lane & 31;
    const int n = __builtin_amdgcn_readfirstlane(*n_list);
    const int64_t total = Npad / CB;
    const bool cn_lane = threadIdx.x < CB / 4;

    for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += (int64_t)gridDim.x * 256) {
        int64_t qrow[2], qr[2];
#pragma unroll
        for (int qb = 0; qb < 2; ++qb) {
            const int64_t e = base + wave * 64 + qb * 32 + col;
            qrow[qb] = (e < n) ? list[e] : -1;
            qr[qb] = qrow[qb] < 0 ? 0 : qrow[qb];
        }
        uint4 qf[2][Sh::KS];
        sym_load_queries<DK>(Xh, qr, h, qf);
        float ld[2][16];
        int li[2][16];
        float thr[2] = {INFINITY, INFINITY};
#pragma unroll
        for (int qb = 0; qb < 2; ++qb)
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                ld[qb][p] = INFINITY;
                li[qb][p] = -1;
            }
        u32x4_t stage[Sh::NV];
        float4 scn = make_float4(0.f, 0.f, 0.f, 0.f);
        sym_load_stage<DK>(Xh, 0, stage);
        if (cn_lane) scn = *(const float4 *)(cn + 4 * threadIdx.x);
        sym_store_stage<DK>(lds[0], stage);
        if (cn_lane) *(float4 *)(&ldc[0][4 * threadIdx.x]) = scn;
        __syncthreads();
        for (int64_t st = 0; st < total; ++st) {
            const int buf = (int)(st & 1);
            const int64_t nx = (st + 1 < total) ? st + 1 : st;
            sym_load_stage<DK>(Xh, nx * CB, stage);
            if (cn_lane) scn = *(const float4 *)(cn + nx * CB + 4 * threadIdx.x);
#pragma unroll
            for (int rb = 0; rb < CB / 32; ++rb) {
                const int64_t c0 = st * CB + rb * 32;
                f32x16_t acc[2];
                sym_mfma<P, DK>(&lds[buf][(rb * 32 + col) * Sh::LROWB + h * 16], qf, acc);
                float cnv[16];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float4 a = *(const float4 *)(&ldc[buf][rb * 32 + 8 * g + 4 * h]);
                    cnv[4 * g] = a.x;
                    cnv[4 * g + 1] = a.y;
                    cnv[4 * g + 2] = a.z;
                    cnv[4 * g + 3] = a.w;
                }
#pragma unroll
                for (int qb = 0; qb < 2; ++qb) {
                    float m = INFINITY;
#pragma unroll
                    for (int i = 0; i < 16; ++i) m = fminf(m, fmaf(beta, acc[qb][i], cnv[i]));
                    if (__any(m < thr[qb])) {
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const float sc = fmaf(beta, acc[qb][i], cnv[i]);
                            const bool pass = sc < thr[qb];
                            if (__any(pass)) {
                                if (pass) {
                                    list_insert<16>(ld[qb], li[qb], sc, (int)(c0 + 8 * (i >> 2) + 4 * h + (i & 3)));
                                    thr[qb] = ld[qb][15];
                                }
                            }
                        }
                    }
                }
            }
            sym_store_stage<DK>(lds[buf ^ 1], stage);
            if (cn_lane) *(float4 *)(&ldc[buf ^ 1][4 * threadIdx.x]) = scn;
            __syncthreads();
        }
#pragma unroll
        for (int qb = 0; qb < 2; ++qb) {
            if (qrow[qb] >= 0) {
                int32_t *o = cand + qrow[qb] * 32 + h * 16;
#pragma unroll
                for (int p = 0; p < 16; ++p) o[p] = li[qb][p];
            }
        }
    }
}

__global__ void k_sym_zero(int *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_sym_fill(float *__restrict__ p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------ host side
constexpr int kSymBucket = 512;  // bucket entries per row

size_t sym_bucket_entries() { return kSymBucket; }

template <int P, int DK>
static void sym_launch(const void *Xh, const float *cn, int64_t N, int64_t Npad, float beta, int K1, int stride,
                       int split, const SymBuffers &w, int32_t *cand, hipStream_t st) {
    const float sigma = (float)(32 * DK) * 2048.0f;  // rounding slack (see knn_sym_enqueue)
    const int B = kSymBucket;
    const int64_t Ns = (N + stride - 1) / stride;
    k_sym_zero<<<cdiv(Npad + 1, 256), 256, 0, st>>>(w.cnt, Npad + 1);  // counters and the fallback count
    NCVIS_LAUNCH_CHECK();
    // half-lists of ceil(K1 / 2): the bound covers the (K1 or K1 + 1)-th smallest score
    if (K1 <= 8)
        k_sym_bound<P, DK, 4><<<(unsigned)(Npad / 256), 256, 0, st>>>(Xh, cn, N, beta, stride, Ns, sigma, w.T);
    else
        k_sym_bound<P, DK, 8><<<(unsigned)(Npad / 256), 256, 0, st>>>(Xh, cn, N, beta, stride, Ns, sigma, w.T);
    NCVIS_LAUNCH_CHECK();
    const int64_t nbI = Npad / 256;
    const int NS = (int)(Npad / SymDma<DK>::CB), SPB = SymDma<DK>::SPB;
    int chunk = tuning().knn_sym_chunk;
    if (chunk > 0) chunk = std::max(SPB, chunk / SPB * SPB);  // whole query blocks per chunk
    if (chunk > 0) {
        const int64_t items = sym_chunk_prefix((NS + chunk - 1) / chunk, chunk / SPB, nbI);
        k_sym_main<P, DK><<<dim3((unsigned)items, 1u), 256, 0, st>>>(Xh, cn, w.T, N, Npad, beta, sigma, B, w.cnt,
                                                                     w.bucket, chunk);
    } else {
        k_sym_main<P, DK><<<dim3((unsigned)((nbI + 1) / 2), (unsigned)split), 256, 0, st>>>(
            Xh, cn, w.T, N, Npad, beta, sigma, B, w.cnt, w.bucket, 0);
    }
    NCVIS_LAUNCH_CHECK();
    int *n_ovf = w.cnt + Npad;
    k_sym_select<<<cdiv(N, 4), 256, 0, st>>>(w.cnt, w.bucket, B, N, K1, cand, w.ovf, n_ovf);
    NCVIS_LAUNCH_CHECK();
    const unsigned G = (unsigned)std::min<int64_t>(512, std::max<int64_t>(1, cdiv(N, 256)));
    k_sym_fallback<P, DK><<<G, 256, 0, st>>>(Xh, cn, Npad, beta, w.ovf, n_ovf, cand);
    NCVIS_LAUNCH_CHECK();
}

// Scores and bounds live on the GEMM input's scale: every transformed element is at most 2^14 in
// magnitude, so cn, beta <q, c> and every bound stay below 4 Dpad 2^28, and each rounding in the
// candidate-side filter is below 64 Dpad; sigma = 2048 Dpad covers them and also keeps a bound
// inclusive when a pair's inner product is rounded differently with the operands' roles exchanged.
void knn_sym_enqueue(int prec, int DK, const void *Xh, const float *cn, int64_t N, int64_t Npad, float beta,
                     int K1, int stride, int split, const SymBuffers &w, int32_t *cand, hipStream_t st) {
    if (Npad % 256 != 0 || K1 < 1 || K1 > 16 || stride < 1 || (N + stride - 1) / stride < 64)
        throw HipError("[ncvis_amd] sym search: bad shape");
#define NCVIS_SYM(PV)                                                                       \
    switch (DK) {                                                                           \
        case 1: sym_launch<PV, 1>(Xh, cn, N, Npad, beta, K1, stride, split, w, cand, st); break; \
        case 2: sym_launch<PV, 2>(Xh, cn, N, Npad, beta, K1, stride, split, w, cand, st); break; \
        case 4: sym_launch<PV, 4>(Xh, cn, N, Npad, beta, K1, stride, split, w, cand, st); break; \
        default: throw HipError("[ncvis_amd] sym search: unsupported padded dimension");     \
    }
    if (prec == PREC_BF16) {
        NCVIS_SYM(PREC_BF16)
    } else if (prec == PREC_FP16) {
        NCVIS_SYM(PREC_FP16)
    } else {
        throw HipError("[ncvis_amd] sym search: 16-bit GEMMs only");
    }
#undef NCVIS_SYM
}

}  // namespace ncvis_hip
