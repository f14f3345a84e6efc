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
//   k_knn_tiles   - MFMA tiles (knn_tiles.h, instantiated in knn_t16 / knn_q1_* / knn_wide): a
//                   workgroup owns 4 waves x 32*QB queries whose B fragments live in registers;
//                   candidate rows stream through double-buffered LDS stages; every 32x32 MFMA
//                   block is scored (cn + beta * <q, c>) and filtered into per-lane sorted top-KL
//                   lists (two half-lists per query, one per lane half)
//                   (FP32 rows of 257..512 elements: 16x16x4 tiles with four quarter-lists per
//                   query, knn_wide_f32.hip)
//   k_knn_rerank  - exact FP32 distances for the 2*KL (4*KL) candidates of each query, rank by
//                   (distance, label), drop rank 0, write ranks k..1 (farthest first)
//   k_knn_generic - the remaining shapes: exact FP32 scan (LDS-tiled) with a max-heap per query
#include <algorithm>
#include <cfloat>
#include <climits>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

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

// ------------------------------------------------------------------------ GEMM input transform
// The candidate GEMM ranks candidates by ||c||^2 - 2<q,c>, computed on rounded copies of the
// rows. Two transforms keep that ranking meaningful for any input scale without changing it:
//   * L2 only: subtract a per-column offset c_j (translation leaves every distance unchanged);
//     without it, data on a large offset cancels catastrophically in the expansion and loses its
//     within-cluster resolution in FP16/BF16. c_j is the column mean rounded to a multiple of
//     2^floor(log2(max_i |x_ij - mean_j|)), a grid coarse enough that integer data stays integer
//     (exact in every GEMM precision) and fine enough to leave at most half a spread of offset.
//   * every distance: one global power-of-two scale s putting max |x - c| in (2^13, 2^14], inside
//     FP16's range (|x| > 65504 would become inf) and above its subnormals. Scaling by a power of
//     two multiplies every score by s^2 exactly, so the order is unchanged.
// Non-finite statistics (non-finite inputs) fall back to c = 0, s = 1.
// The rerank recomputes exact distances on the untransformed rows.

// Per-column partial statistics over a strided range of rows: MODE 0 sums x (FP64); MODE 1 the
// maximum of |x - mean_j| (mean may be null: max |x|). Block (64, 4), grid (ceil(D / 64), nb).
template <int MODE>
__global__ __launch_bounds__(256) void k_colscan(const float *__restrict__ X, int64_t N, int D,
                                                 const double *__restrict__ mean, double *__restrict__ part) {
    __shared__ double red[4][64];
    const int j = blockIdx.x * 64 + threadIdx.x;
    double acc = 0.0;
    if (j < D) {
        const double m = (MODE == 1 && mean) ? mean[j] : 0.0;
        for (int64_t i = (int64_t)blockIdx.y * 4 + threadIdx.y; i < N; i += (int64_t)gridDim.y * 4) {
            const double v = (double)X[i * D + j];
            if (MODE == 0)
                acc += v;
            else
                acc = fmax(acc, fabs(v - m)) + (v - v);  // + (v - v): a NaN input propagates
        }
    }
    red[threadIdx.y][threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.y == 0 && j < D) {
        double r = red[0][threadIdx.x];
        for (int t = 1; t < 4; ++t) {
            const double v = red[t][threadIdx.x];
            r = (MODE == 0) ? r + v : fmax(r, v) + (v - v);
        }
        part[(int64_t)blockIdx.y * D + j] = r;
    }
}

// mean[j] = sum of the nb partial sums / N (one thread per column)
__global__ void k_colmean(const double *__restrict__ part, int nb, int64_t N, int D, double *__restrict__ mean) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= D) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(int64_t)b * D + j];
    mean[j] = s / (double)N;
}

// One block: cs[j] = c_j (0 without centring), cs[D] = s, from the means (null: no centring)
// and the nb partial maxima of |x - mean_j|.
__global__ __launch_bounds__(256) void k_pack_params(const double *__restrict__ mean, const double *__restrict__ pdev,
                                                     int nb, int D, float *__restrict__ cs) {
    __shared__ double red_m[256];
    __shared__ int red_bad[256];
    double M = 0.0;
    int bad = 0;
    for (int j = threadIdx.x; j < D; j += 256) {
        double dev = 0.0;
        for (int b = 0; b < nb; ++b) {
            const double v = pdev[(int64_t)b * D + j];
            dev = fmax(dev, v) + (v - v);
        }
        const double mu = mean ? mean[j] : 0.0;
        double c = 0.0;
        if (mean) c = (dev > 0.0) ? rint(mu / exp2(floor(log2(dev)))) * exp2(floor(log2(dev))) : mu;
        const float cf = (float)c;
        const double m = dev + fabs(mu - (double)cf);
        if (!isfinite(m)) bad = 1;
        M = fmax(M, m);
        cs[j] = cf;
    }
    red_m[threadIdx.x] = M;
    red_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red_m[threadIdx.x] = fmax(red_m[threadIdx.x], red_m[threadIdx.x + w]);
            red_bad[threadIdx.x] |= red_bad[threadIdx.x + w];
        }
        __syncthreads();
    }
    bad = red_bad[0];
    M = red_m[0];
    if (bad)
        for (int j = threadIdx.x; j < D; j += 256) cs[j] = 0.f;
    if (threadIdx.x == 0) {
        float sc = 1.f;
        if (!bad && M > 0.0) {
            int e = 14 - (int)ceil(log2(M));
            e = e < -100 ? -100 : (e > 100 ? 100 : e);
            sc = (float)exp2((double)e);
        }
        cs[D] = sc;
    }
}

// rows per k_pack block: LDS holds R rows of Dpad + 1 floats (at most 33 KB)
inline int pack_rows(int Dpad) { return Dpad <= 256 ? 32 : 16; }

// Element k of a padded row as stored for MFMA: (x_k - c_k) * s, rounded to the GEMM precision.
// 16-bit precisions keep the natural order (lane half h of a 32x32x16 step reads
// k = 16 s + 8 h .. +7, contiguous). FP32 is stored permuted so a lane's operands are contiguous
// too: with G lane groups splitting each step (32x32x2: G = 2, group h reads k = 2 s + h;
// 16x16x4: G = 4, group g reads k = 4 s + g) element k = G s + g sits at position g (Dpad / G) + s.
// A and B use the same permutation, so dot products are unchanged.
template <int P, int G = 2>
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ Xp, int64_t N, int64_t Npad, int D, int Dpad,
                                              int is_l2, const float *__restrict__ cs, void *__restrict__ Xh,
                                              float *__restrict__ cn, int R) {
    // R rows per block: the 256 threads read the rows' elements coalesced, write the
    // rounded elements and keep them in LDS in stored order; one thread per row then sums their
    // squares sequentially in stored order (the score offset cn)
    extern __shared__ float lds_pack[];  // [R][Dpad + 1]
    const int64_t r0 = (int64_t)blockIdx.x * R;
    const float sc = cs[D];
    const int part = Dpad / G;
    const int pitch = Dpad + 1;
    for (int e = threadIdx.x; e < R * Dpad; e += 256) {
        const int row = e / Dpad, kk = e - row * Dpad;  // kk: element index (natural order)
        const int64_t i = r0 + row;
        if (i >= Npad) continue;
        const float v = (i < N && kk < D) ? __fmul_rn(__fsub_rn(Xp[i * D + kk], cs[kk]), sc) : 0.f;
        float r;
        int pos;
        if (P == PREC_FP32) {
            pos = (kk % G) * part + kk / G;  // element G st + g sits at position g (Dpad / G) + st
            ((float *)Xh)[i * Dpad + pos] = v;
            r = v;
        } else {
            pos = kk;
            uint16_t bits;
            if (P == PREC_BF16) {
                bits = f32_to_bf16_rne(v);
                r = bf16_to_f32(bits);
            } else {
                const _Float16 hv = (_Float16)v;
                bits = __builtin_bit_cast(uint16_t, hv);
                r = (float)hv;
            }
            ((uint16_t *)Xh)[i * Dpad + pos] = bits;
        }
        lds_pack[row * pitch + pos] = r;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int64_t i = r0 + threadIdx.x;
        if (i < Npad) {
            float s = 0.f;
            for (int p = 0; p < Dpad; ++p) {
                const float r = lds_pack[threadIdx.x * pitch + p];
                s = __fadd_rn(s, __fmul_rn(r, r));
            }
            cn[i] = (i < N) ? (is_l2 ? s : 0.f) : INFINITY;
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
// (same value as exact_dist). Rows must be 16-byte aligned (D % 4 == 0 and an aligned base). The
// loads go out a 128-byte line at a time (8 float4 of each row) before their terms are summed, so
// a lane's requests for one line arrive together instead of one per step of the sum.
__device__ inline void dist4_terms(float &s, float4 a, float4 b, int is_l2) {
    if (is_l2) {
        float t = __fsub_rn(a.x, b.x);
        s = __fadd_rn(s, __fmul_rn(t, t));
        t = __fsub_rn(a.y, b.y);
        s = __fadd_rn(s, __fmul_rn(t, t));
        t = __fsub_rn(a.z, b.z);
        s = __fadd_rn(s, __fmul_rn(t, t));
        t = __fsub_rn(a.w, b.w);
        s = __fadd_rn(s, __fmul_rn(t, t));
    } else {
        s = __fadd_rn(s, __fmul_rn(a.x, b.x));
        s = __fadd_rn(s, __fmul_rn(a.y, b.y));
        s = __fadd_rn(s, __fmul_rn(a.z, b.z));
        s = __fadd_rn(s, __fmul_rn(a.w, b.w));
    }
}

__device__ inline float exact_dist4(const float *__restrict__ x, const float *__restrict__ y, int D, int is_l2) {
    const float4 *x4 = (const float4 *)x;
    const float4 *y4 = (const float4 *)y;
    const int n4 = D / 4;
    float s = 0.f;
    int j = 0;
    for (; j + 8 <= n4; j += 8) {
        float4 a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = x4[j + u];
            b[u] = y4[j + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) dist4_terms(s, a[u], b[u], is_l2);
    }
    for (; j < n4; ++j) dist4_terms(s, x4[j], y4[j], is_l2);
    return is_l2 ? s : __fsub_rn(1.0f, s);
}

// C = 2*KL candidates per query (32 or 64); a wave reranks 64/C queries, lane c of a query's
// segment handles candidate c: exact FP32 distance, rank by (distance, label, slot) with
// shuffles inside the segment, drop rank 0, write ranks k..1 (farthest first).
template <int C>
__global__ __launch_bounds__(256) void k_knn_rerank(const float *__restrict__ Xp, int64_t N, int D, int is_l2,
                                                    int vec4, const int32_t *__restrict__ cand, int k,
                                                    int32_t *__restrict__ inds, float *__restrict__ dists,
                                                    int *__restrict__ err, int32_t label_off) {
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
        if (ci == INT_MAX) atomicOr(err, kErrKnnCandidates);
        inds[q * k + (k - rank)] = ((ci == INT_MAX) ? 0 : ci) + label_off;
        dists[q * k + (k - rank)] = dv;
    }
}

// 2*KL = 128 candidates per query (k + 1 up to 64): a wave per query, lanes l and l + 64 of the
// candidate list on each lane; same ranking and output as k_knn_rerank.
__global__ __launch_bounds__(256) void k_knn_rerank128(const float *__restrict__ Xp, int64_t N, int D, int is_l2,
                                                       int vec4, const int32_t *__restrict__ cand, int k,
                                                       int32_t *__restrict__ inds, float *__restrict__ dists,
                                                       int *__restrict__ err, int32_t label_off) {
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
            if (ci[u] == INT_MAX) atomicOr(err, kErrKnnCandidates);
            inds[q * k + (k - rank[u])] = ((ci[u] == INT_MAX) ? 0 : ci[u]) + label_off;
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
                                                     int32_t *__restrict__ inds, float *__restrict__ dists,
                                                     int32_t label_off) {
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
        inds[q * k + j] = hi[K1 - 1 - j] + label_off;
        dists[q * k + j] = hd[K1 - 1 - j];
    }
}

// The shape decisions of one search and the carving of its workspace (256-byte boundaries).
struct KnnLayout {
    bool tiled, wide, wide32;
    int DK, Dpad, QB, KL, lists, C, esz, nb;
    int64_t Npad;
    TileVariant tv;
    size_t xp, part, mean, cs, xh, cn, cand, hd, hi, sT, scnt, sbucket, sovf, total;
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Largest candidate split a search admits: every range keeps at least 128 candidate rows (the
// largest LDS stage) and the rerank takes at most 128 candidates per query.
static int max_split(int64_t Npad, int lists, int KL) {
    int s = 1;
    while (s < 4 && Npad / (2 * s) >= 128 && lists * KL * 2 * s <= 128) s *= 2;
    return s;
}

static KnnLayout knn_layout(int64_t N, int D, int k, int dist, int prec, const TileVariant *force = nullptr) {
    KnnLayout L{};
    // rows up to 256 elements: 32x32 MFMA tiles in every precision (k + 1 <= 64); rows of
    // 257..512: one query set per wave in the 16-bit GEMMs, 16x16x4 tiles in FP32 (k + 1 <= 32);
    // wider rows or longer lists take the generic exact scan
    L.wide = (D > 256) && (D <= 512);
    L.wide32 = L.wide && prec == PREC_FP32;
    L.KL = (k + 1 <= 16) ? 16 : (k + 1 <= 32) ? 32 : 64;
    L.tiled = (D <= 256 && k + 1 <= 64) || (L.wide && k + 1 <= 32);
    L.DK = D <= 32 ? 1 : D <= 64 ? 2 : D <= 128 ? 4 : D <= 256 ? 8 : D <= 320 ? 10 : D <= 384 ? 12 : 16;
    L.Dpad = 32 * L.DK;
    L.QB = (L.wide || L.KL >= 32 || prec == PREC_FP32) ? 1 : 2;
    const int64_t qtile = L.wide32 ? 64 : 4 * 32 * L.QB;
    L.Npad = ((N + qtile - 1) / qtile) * qtile;
    L.lists = L.wide32 ? 4 : 2;  // candidate lists per query and candidate range
    if (L.tiled) {
        const int smax = max_split(L.Npad, L.lists, L.KL);
        // symmetric threshold search (knn_sym.hip): 16-bit GEMMs, k + 1 <= 16, rows of at most
        // 128 elements, at least two query blocks and 64 bounding rows, and at most 2^23 rows (its
        // buckets take 4 KiB per row)
        const bool sym_ok = !L.wide && L.KL == 16 && L.DK <= 4 && prec != PREC_FP32 && L.Npad >= 512 &&
                            N <= ((int64_t)1 << 23);
        if (force) {
            L.tv = *force;
        } else if (!tuned_variant(prec, L.DK, L.KL, L.wide, N, L.tv)) {
            const Tuning &tu = tuning();
            L.tv.stage_mult = tu.knn_stage_mult;
            L.tv.dma = tu.knn_dma != 0 ? 1 : 0;
            L.tv.sym = (tu.knn_sym >= 0) ? tu.knn_sym : (N >= ((int64_t)1 << 18) ? 8 : 0);
            if (!sym_ok) L.tv.sym = 0;
            L.tv.split = 1;
            if (tu.knn_split > 0) {
                L.tv.split = tu.knn_split;
            } else if (L.tv.sym > 0) {
                // the main pass pairs query blocks: split their sweeps until about 2048 workgroups
                const int64_t pairs = (L.Npad / 256 + 1) / 2;
                while (L.tv.split < 4 && pairs * L.tv.split < 2048) L.tv.split *= 2;
            } else {
                // list tiles: split the candidates until the launch has about 1024 workgroups
                const int64_t blocks = L.Npad / qtile;
                while (L.tv.split < smax && blocks * L.tv.split < 1024) L.tv.split *= 2;
            }
        }
        if (!sym_ok || L.tv.sym < 0 || (N + std::max(L.tv.sym, 1) - 1) / std::max(L.tv.sym, 1) < 64) L.tv.sym = 0;
        if (L.tv.split < 1) L.tv.split = 1;
        if (L.tv.split > (L.tv.sym > 0 ? 4 : smax)) L.tv.split = (L.tv.sym > 0) ? 4 : smax;
    }
    // candidates per query for the rerank: 32, 64 or 128 (the symmetric search writes 32)
    L.C = (L.tv.sym > 0) ? 32 : L.lists * L.KL * L.tv.split;
    L.esz = (prec == PREC_FP32) ? 4 : 2;
    L.nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, cdiv(N, 256)));
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.xp = take((dist == DIST_COSINE || dist == DIST_CORRELATION) ? sizeof(float) * N * D : 0);
    if (L.tiled) {
        L.part = take(sizeof(double) * L.nb * D);
        L.mean = take(sizeof(double) * D);
        L.cs = take(sizeof(float) * (D + 1));
        L.xh = take((size_t)L.Npad * L.Dpad * L.esz);
        L.cn = take(sizeof(float) * L.Npad);
        L.cand = take(sizeof(int32_t) * N * L.C);
        if (L.tv.sym > 0) {
            L.sT = take(sizeof(float) * L.Npad);
            L.scnt = take(sizeof(int) * (L.Npad + 1));
            L.sbucket = take(sizeof(uint2) * N * sym_bucket_entries());
            L.sovf = take(sizeof(int) * N);
        }
    } else {
        L.hd = take(sizeof(float) * N * (k + 1));
        L.hi = take(sizeof(int) * N * (k + 1));
    }
    L.total = std::max<size_t>(o, 256);
    return L;
}

size_t knn_workspace_bytes(int64_t N, int D, int k, int dist, int prec) {
    return knn_layout(N, D, k, dist, prec).total;
}

TileVariant knn_variant(int64_t N, int D, int k, int dist, int prec) { return knn_layout(N, D, k, dist, prec).tv; }

size_t knn_workspace_bytes_for(int64_t N, int D, int k, int dist, int prec, const TileVariant &tv) {
    return knn_layout(N, D, k, dist, prec, &tv).total;
}

TileVariant knn_effective_variant(int64_t N, int D, int k, int dist, int prec, const TileVariant &tv) {
    return knn_layout(N, D, k, dist, prec, &tv).tv;
}

// cs [D + 1]: per-column offsets (0 unless centring) and the scale, all on the device.
static void gemm_input_transform(const float *Xp, int64_t N, int D, int center, int nb, double *part, double *mean,
                                 float *cs, hipStream_t st) {
    const dim3 grid((unsigned)cdiv(D, 64), (unsigned)nb), block(64, 4);
    if (center) {
        k_colscan<0><<<grid, block, 0, st>>>(Xp, N, D, nullptr, part);
        k_colmean<<<cdiv(D, 256), 256, 0, st>>>(part, nb, N, D, mean);
        NCVIS_LAUNCH_CHECK();
    }
    k_colscan<1><<<grid, block, 0, st>>>(Xp, N, D, center ? mean : nullptr, part);
    k_pack_params<<<1, 256, 0, st>>>(center ? mean : nullptr, part, nb, D, cs);
    NCVIS_LAUNCH_CHECK();
}

void knn_enqueue(const float *dX, int64_t N, int D, int k, int dist, int prec, int32_t *d_inds, float *d_dists,
                 void *ws, int *d_err, hipStream_t st, const KnnEvents *ev, const TileVariant *force,
                 int32_t label_off) {
    if (N < 2 || D < 1 || k < 1 || k > N - 1) throw HipError("[ncvis_amd] knn: invalid sizes");
    if (N > (int64_t)INT_MAX - 512) throw HipError("[ncvis_amd] knn: more than 2^31 points");
    if (prec < PREC_FP32 || prec > PREC_FP16) throw HipError("[ncvis_amd] knn: unknown precision");
    const KnnLayout L = knn_layout(N, D, k, dist, prec, force);
    const int S = L.tv.split;
    unsigned char *w = (unsigned char *)ws;
    const int is_l2 = (dist == DIST_L2);
    const float *Xp = dX;
    if (dist == DIST_COSINE || dist == DIST_CORRELATION) {
        float *xp = (float *)(w + L.xp);
        k_preprocess<<<cdiv(N, 256), 256, 0, st>>>(dX, N, D, dist, xp);
        NCVIS_LAUNCH_CHECK();
        Xp = xp;
    }
    if (L.tiled) {
        const int DK = L.DK, Dpad = L.Dpad, KL = L.KL;
        const int64_t Npad = L.Npad;
        void *xh = w + L.xh;
        float *cn = (float *)(w + L.cn);
        float *cs = (float *)(w + L.cs);
        int32_t *cand = (int32_t *)(w + L.cand);
        gemm_input_transform(Xp, N, D, is_l2, L.nb, (double *)(w + L.part), (double *)(w + L.mean), cs, st);
        const int pr = pack_rows(Dpad);
        const unsigned pb = (unsigned)cdiv(Npad, pr);
        const size_t plds = sizeof(float) * pr * (Dpad + 1);
        switch (prec) {
            case PREC_FP32:
                if (L.wide32)
                    k_pack<PREC_FP32, 4><<<pb, 256, plds, st>>>(Xp, N, Npad, D, Dpad, is_l2, cs, xh, cn, pr);
                else
                    k_pack<PREC_FP32, 2><<<pb, 256, plds, st>>>(Xp, N, Npad, D, Dpad, is_l2, cs, xh, cn, pr);
                break;
            case PREC_BF16:
                k_pack<PREC_BF16><<<pb, 256, plds, st>>>(Xp, N, Npad, D, Dpad, is_l2, cs, xh, cn, pr);
                break;
            default:
                k_pack<PREC_FP16><<<pb, 256, plds, st>>>(Xp, N, Npad, D, Dpad, is_l2, cs, xh, cn, pr);
                break;
        }
        NCVIS_LAUNCH_CHECK();
        const float beta = is_l2 ? -2.0f : -1.0f;
        if (ev) record_event(ev->tiles_begin, st);
        if (L.wide32)
            knn_tiles_wide_f32(KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        else if (L.wide)
            knn_tiles_wide(prec, KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        else if (L.tv.sym > 0)
            knn_sym_enqueue(prec, DK, xh, cn, N, Npad, beta, k + 1, L.tv.sym, S,
                            SymBuffers{(float *)(w + L.sT), (int *)(w + L.scnt), (uint2 *)(w + L.sbucket),
                                       (int *)(w + L.sovf)},
                            cand, st);
        else if (KL == 16)
            knn_tiles16(prec, DK, N, Npad, xh, cn, beta, cand, st, L.tv);
        else if (prec == PREC_FP32 && L.tv.dma != 0)
            knn_tiles_q1_f32dma(KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        else if (prec == PREC_FP32)
            knn_tiles_q1_f32(KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        else if (prec == PREC_BF16)
            knn_tiles_q1_bf16(KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        else
            knn_tiles_q1_fp16(KL, DK, N, Npad, xh, cn, beta, cand, st, S);
        NCVIS_LAUNCH_CHECK();
        if (ev) record_event(ev->tiles_end, st);
        const int vec4 = (D % 4 == 0) && ((uintptr_t)Xp % 16 == 0);
        if (L.C == 32)
            k_knn_rerank<32><<<cdiv(N, 8), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand, k, d_inds, d_dists, d_err,
                                                            label_off);
        else if (L.C == 64)
            k_knn_rerank<64><<<cdiv(N, 4), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand, k, d_inds, d_dists, d_err,
                                                            label_off);
        else
            k_knn_rerank128<<<cdiv(N, 4), 256, 0, st>>>(Xp, N, D, is_l2, vec4, cand, k, d_inds, d_dists, d_err,
                                                            label_off);
        NCVIS_LAUNCH_CHECK();
        if (ev) record_event(ev->rerank_end, st);
    } else {
        if (ev) record_event(ev->tiles_begin, st);
        k_knn_generic<<<cdiv(N, 256), 256, 0, st>>>(Xp, N, D, is_l2, k, (float *)(w + L.hd), (int *)(w + L.hi),
                                                    d_inds, d_dists, label_off);
        NCVIS_LAUNCH_CHECK();
        if (ev) {
            record_event(ev->tiles_end, st);
            record_event(ev->rerank_end, st);
        }
    }
}

__global__ void k_zero_int(int *p) { *p = 0; }

void knn_run(const float *dX, int64_t N, int D, int k, int dist, int prec, int32_t *d_inds, float *d_dists,
             hipStream_t st, KnnStats *stats) {
    if (N < 2 || D < 1 || k < 1 || k > N - 1) throw HipError("[ncvis_amd] knn: invalid sizes");
    DevBuf<unsigned char> ws(knn_workspace_bytes(N, D, k, dist, prec));
    DevBuf<int> err(1);
    k_zero_int<<<1, 1, 0, st>>>(err.p);
    KnnEvents ev{};
    NCVIS_HIP_CHECK(hipEventCreate(&ev.tiles_begin));
    NCVIS_HIP_CHECK(hipEventCreate(&ev.tiles_end));
    NCVIS_HIP_CHECK(hipEventCreate(&ev.rerank_end));
    struct EvGuard {
        KnnEvents &e;
        ~EvGuard() {
            (void)hipEventDestroy(e.tiles_begin);
            (void)hipEventDestroy(e.tiles_end);
            (void)hipEventDestroy(e.rerank_end);
        }
    } evg{ev};
    knn_enqueue(dX, N, D, k, dist, prec, d_inds, d_dists, ws.p, err.p, st, &ev);
    int herr = 0;
    NCVIS_HIP_CHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (stats) {
        float t1 = 0, t2 = 0;
        NCVIS_HIP_CHECK(hipEventElapsedTime(&t1, ev.tiles_begin, ev.tiles_end));
        NCVIS_HIP_CHECK(hipEventElapsedTime(&t2, ev.tiles_end, ev.rerank_end));
        stats->tile_ms = t1;
        stats->rerank_ms = t2;
        stats->used_mfma = knn_layout(N, D, k, dist, prec).tiled ? 1 : 0;
    }
    if (herr) throw HipError("[ncvis_amd] knn: fewer than k+1 finite candidates (non-finite input?)");
}

}  // namespace ncvis_hip
