// Layout initialisation, NCE SGD epochs and the explicit NCE loss / adjoint on gfx950.
//
// Initialisation (src/ncvis.cpp:184-273):
//   k_init_random  Y[i] = uniform_real<float>(0,1) drawn from pcg64(seed + t), where t is the
//                  thread that owns element i under libgomp's static schedule of the N*d loop
//                  (ncvis.cpp:198-201); a GPU thread jumps the stream to its first element.
//   k_spmm_rows    Y_new[i] = sum_{edges (i, j) in row order} alpha * Y_old[j], row-owned
//                  (race-free), uncontracted like the reference's loop (ncvis.cpp:206-218)
//   k_colsum / k_colvar / k_finalize / k_normalize
//                  per-dimension mean and population std (FP64 partial sums, deterministic)
//                  and standardisation (ncvis.cpp:220-260)
// Epochs (src/ncvis.cpp:275-352):
//   k_sgd_epoch    T chains, chain t = the reference's OpenMP thread t: pcg64(seed + t), the
//                  static edge chunk [s0, e0) of thread t, a private Q copy, sequential samples.
//                  Chains run concurrently Hogwild-style; T = 1 is the reference at one thread.
//   k_sgd_epoch_hw the many-chain form: source row kept in registers with the chain's own
//                  updates, other endpoints updated atomically (one 64-bit compare-and-swap per
//                  coordinate pair), negatives' rows prefetched.
//   k_q_reduce     Q = mean of the chains' Q copies (ncvis.cpp:342-349), FP64 sum.
// NCE loss (the objective the update of ncvis.cpp:304-338 descends):
//   k_nce_loss     per edge: positive + n explicit negatives; loss and dL/dQ as FP64 partial
//                  sums, dL/dY scattered with FP32 atomics.
#include <algorithm>
#include <cfloat>
#include <stdexcept>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

// ------------------------------------------------------------------------ init
__global__ void k_init_random(float *__restrict__ Y, int64_t Nd, int64_t T, uint64_t seed) {
    constexpr int RUN = 64;
    int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * RUN;
    if (i0 >= Nd) return;
    int64_t i1 = i0 + RUN < Nd ? i0 + RUN : Nd;
    // owner of element i0 under the static schedule
    int64_t q = Nd / T, r = Nd % T, t;
    if (i0 < r * (q + 1))
        t = i0 / (q + 1);
    else
        t = r + (i0 - r * (q + 1)) / q;
    int64_t s0, e0;
    static_chunk(Nd, T, t, s0, e0);
    U128 g = pcg_advance(pcg_seed(seed + (uint64_t)t), (uint64_t)(i0 - s0));
    for (int64_t i = i0; i < i1; ++i) {
        while (i >= e0) {
            ++t;
            static_chunk(Nd, T, t, s0, e0);
            g = pcg_seed(seed + (uint64_t)t);
        }
        Y[i] = uniform_unit_float(g);
    }
}

__global__ void k_spmm_rows(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col, int64_t N, int d,
                            float alpha, const float *__restrict__ Yo, float *__restrict__ Yn) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t b = row_ptr[i], e = row_ptr[i + 1];
    for (int c = 0; c < d; ++c) {
        float acc = 0.f;
        for (int64_t p = b; p < e; ++p) acc = __fadd_rn(acc, __fmul_rn(alpha, Yo[(int64_t)col[p] * d + c]));
        Yn[i * d + c] = acc;
    }
}

// Same sums for d <= 4 with the dimension loop inside the edge loop: each edge's column index
// and row are read once, four edges are fetched ahead of their (in-order) accumulation. Every
// dimension's sum keeps the sequential edge order, so the result equals k_spmm_rows bit for bit.
template <int DIM>
__global__ void k_spmm_rows_d(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col, int64_t N,
                              float alpha, const float *__restrict__ Yo, float *__restrict__ Yn) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t b = row_ptr[i], e = row_ptr[i + 1];
    float acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = 0.f;
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

// per-block FP64 column sums of Y (mode 0) or of (Y - mean)^2 (mode 1)
template <int MODE>
__global__ __launch_bounds__(256) void k_colstat(const float *__restrict__ Y, int64_t N, int d,
                                                 const float *__restrict__ mean, double *__restrict__ partial) {
    __shared__ double red[256];
    for (int c = 0; c < d; ++c) {
        double s = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
            float v = Y[i * d + c];
            if (MODE == 1) {
                float dl = __fsub_rn(v, mean[c]);
                v = __fmul_rn(dl, dl);
            }
            s += (double)v;
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * d + c] = red[0];
        __syncthreads();
    }
}

// MODE 0: mean[c] = sum / N; MODE 1: sigma[c] = sqrtf(sum / N)
template <int MODE>
__global__ void k_finalize(const double *__restrict__ partial, int nb, int d, int64_t N, float *__restrict__ out) {
    int c = threadIdx.x;
    if (c >= d) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * d + c];
    float v = (float)(s / (double)N);
    out[c] = (MODE == 0) ? v : sqrtf(v);
}

__global__ void k_normalize(float *__restrict__ Y, int64_t N, int d, const float *__restrict__ mean,
                            const float *__restrict__ sigma) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * d) return;
    int c = (int)(i % d);
    Y[i] = __fsub_rn(Y[i], mean[c]) / sigma[c];
}

void init_embedding_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, float alpha,
                        int n_init_epochs, uint64_t seed, int64_t chains, float *dY, hipStream_t st) {
    if (N < 1 || d < 1 || d > 1024) throw HipError("[ncvis_amd] init: invalid sizes");
    const int64_t Nd = N * d;
    const int64_t T = chains < 1 ? 1 : chains;
    k_init_random<<<cdiv(cdiv(Nd, 64), 256), 256, 0, st>>>(dY, Nd, T, seed);
    NCVIS_LAUNCH_CHECK();
    if (n_init_epochs <= 0) return;
    DevBuf<float> y2((size_t)Nd);
    const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(N, 256)));
    DevBuf<double> partial((size_t)nb * d);
    DevBuf<float> stats((size_t)2 * d);
    float *Ys[2] = {dY, y2.p};
    for (int ep = 0; ep < n_init_epochs; ++ep) {
        const float *Yo = Ys[ep % 2];
        float *Yn = Ys[(ep + 1) % 2];
        switch (d) {
            case 1: k_spmm_rows_d<1><<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, d_col, N, alpha, Yo, Yn); break;
            case 2: k_spmm_rows_d<2><<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, d_col, N, alpha, Yo, Yn); break;
            case 3: k_spmm_rows_d<3><<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, d_col, N, alpha, Yo, Yn); break;
            case 4: k_spmm_rows_d<4><<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, d_col, N, alpha, Yo, Yn); break;
            default: k_spmm_rows<<<cdiv(N, 256), 256, 0, st>>>(d_row_ptr, d_col, N, d, alpha, Yo, Yn); break;
        }
        NCVIS_LAUNCH_CHECK();
        k_colstat<0><<<nb, 256, 0, st>>>(Yn, N, d, nullptr, partial.p);
        k_finalize<0><<<1, 1024, 0, st>>>(partial.p, nb, d, N, stats.p);
        k_colstat<1><<<nb, 256, 0, st>>>(Yn, N, d, stats.p, partial.p);
        k_finalize<1><<<1, 1024, 0, st>>>(partial.p, nb, d, N, stats.p + d);
        k_normalize<<<cdiv(Nd, 256), 256, 0, st>>>(Yn, N, d, stats.p, stats.p + d);
        NCVIS_LAUNCH_CHECK();
    }
    if (n_init_epochs % 2)
        NCVIS_HIP_CHECK(hipMemcpyAsync(dY, y2.p, sizeof(float) * Nd, hipMemcpyDeviceToDevice, st));
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
    int fast_math;     // 1: single-precision powf / expf
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

// chain_row[t] = row holding edge s0 of chain t (row_ptr[r] <= s0 < row_ptr[r+1])
__global__ void k_chain_rows(const int64_t *__restrict__ row_ptr, int64_t N, int64_t E, int64_t T,
                             int64_t *__restrict__ chain_row, U128 *__restrict__ rng, uint64_t seed) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t s0, e0;
    static_chunk(E, T, t, s0, e0);
    int64_t lo = 0, hi = N;  // largest r with row_ptr[r] <= s0
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) / 2;
        if (row_ptr[mid] <= s0)
            lo = mid;
        else
            hi = mid;
    }
    while (lo < N - 1 && row_ptr[lo + 1] <= s0) ++lo;
    chain_row[t] = lo;
    rng[t] = pcg_seed(seed + (uint64_t)t);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_sgd_epoch(float *__restrict__ Y, int d_rt, const int64_t *__restrict__ row_ptr,
                                                   const int32_t *__restrict__ col, int64_t E, int64_t T,
                                                   const int64_t *__restrict__ chain_row, U128 *__restrict__ rng,
                                                   const float *__restrict__ Qdev, float *__restrict__ qpart,
                                                   SampleConst c, uint64_t N, int atomic_mode) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t s0, e0;
    static_chunk(E, T, t, s0, e0);
    float Qc = *Qdev;
    U128 g = rng[t];
    int64_t r = chain_row[t];
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
    rng[t] = g;
    qpart[t] = Qc;
}

// Many-chain (Hogwild) epoch. A chain keeps the row of its current source point in registers
// and adds its own updates of that row there (the chain sees its own writes, as a thread of the
// reference does), flushing them to Y with atomics when it moves to the next row; updates of
// the other endpoint go to Y with atomics at once, so concurrent chains lose no update. The
// negatives of an edge are drawn before their rows are read (the draw sequence does not depend
// on Y), so PF rows are fetched together instead of one dependent load per sample.
template <int DIM, int PF>
__global__ __launch_bounds__(256) void k_sgd_epoch_hw(float *__restrict__ Y, const int64_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col, int64_t E, int64_t T,
                                                      const int64_t *__restrict__ chain_row, U128 *__restrict__ rng,
                                                      const float *__restrict__ Qdev, float *__restrict__ qpart,
                                                      SampleConst c, uint64_t N, int plain_far) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    int64_t s0, e0;
    static_chunk(E, T, t, s0, e0);
    float Qc = *Qdev;
    U128 g = rng[t];
    int64_t r = chain_row[t];
    int64_t rend = row_ptr[r + 1];
    float base[DIM], acc[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        base[k] = Y[r * DIM + k];
        acc[k] = 0.f;
    }
    const uint32_t total = c.noise + 1;
    for (int64_t e = s0; e < e0; ++e) {
        while (e >= rend) {
#pragma unroll
            for (int k = 0; k < DIM; ++k)
                if (acc[k] != 0.f) __hip_atomic_fetch_add(&Y[r * DIM + k], acc[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ++r;
            rend = row_ptr[r + 1];
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                base[k] = Y[r * DIM + k];
                acc[k] = 0.f;
            }
        }
        const int64_t id = r;
        for (uint32_t j0 = 0; j0 < total; j0 += PF) {
            int64_t oth[PF];
            float yo[PF][DIM];
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                const uint32_t j = j0 + p;
                oth[p] = id;
                if (j < total) {
                    if (j == 0) {
                        oth[p] = col[e];
                    } else {
                        int64_t o;
                        do {
                            o = (int64_t)uniform_index(g, N);
                        } while (o == id);
                        oth[p] = o;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < PF; ++p) {
#pragma unroll
                for (int k = 0; k < DIM; ++k) yo[p][k] = Y[oth[p] * DIM + k];
            }
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                if (j0 + p < total) {
                    float yi[DIM], dx[DIM];
#pragma unroll
                    for (int k = 0; k < DIM; ++k) yi[k] = __fadd_rn(base[k], acc[k]);
                    sgd_math<DIM>(yi, yo[p], (j0 + p) == 0, c, Qc, dx);
#pragma unroll
                    for (int k = 0; k < DIM; ++k) acc[k] = __fadd_rn(acc[k], dx[k]);
                    if (plain_far) {
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
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k)
        if (acc[k] != 0.f) __hip_atomic_fetch_add(&Y[r * DIM + k], acc[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    rng[t] = g;
    qpart[t] = Qc;
}

// One block of 1024 threads; each thread keeps four loads in flight so the 1 MB of chain copies at
// 2^18 chains is read at bandwidth instead of one dependent load per thread per step.
__global__ __launch_bounds__(1024) void k_q_reduce(const float *__restrict__ qpart, int64_t T,
                                                   float *__restrict__ Qdev, float *__restrict__ trace, int epoch) {
    __shared__ double red[1024 / 64];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t t = threadIdx.x;
    for (; t + 3 * 1024 < T; t += 4 * 1024) {
        s0 += (double)qpart[t];
        s1 += (double)qpart[t + 1024];
        s2 += (double)qpart[t + 2 * 1024];
        s3 += (double)qpart[t + 3 * 1024];
    }
    for (; t < T; t += 1024) s0 += (double)qpart[t];
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < 1024 / 64; ++w) tot += red[w];
        float q = (T == 1) ? qpart[0] : (float)(tot / (double)T);
        *Qdev = q;
        trace[epoch] = q;
    }
}

template <int PF>
static void launch_hw(int d, dim3 grid, float *Y, const int64_t *row_ptr, const int32_t *col, int64_t E, int64_t T,
                      const int64_t *chain_row, U128 *rng, const float *Qdev, float *qpart, const SampleConst &c,
                      uint64_t N, int plain_far, hipStream_t st) {
    switch (d) {
        case 1: k_sgd_epoch_hw<1, PF><<<grid, 256, 0, st>>>(Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, plain_far); break;
        case 2: k_sgd_epoch_hw<2, PF><<<grid, 256, 0, st>>>(Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, plain_far); break;
        case 3: k_sgd_epoch_hw<3, PF><<<grid, 256, 0, st>>>(Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, plain_far); break;
        default: k_sgd_epoch_hw<4, PF><<<grid, 256, 0, st>>>(Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, plain_far); break;
    }
}

static void launch_epoch(int d, dim3 grid, float *Y, const int64_t *row_ptr, const int32_t *col, int64_t E,
                         int64_t T, const int64_t *chain_row, U128 *rng, const float *Qdev, float *qpart,
                         const SampleConst &c, uint64_t N, int atomic_mode, hipStream_t st) {
    // the many-chain kernel updates coordinate pairs with 64-bit compare-and-swap, which needs
    // 8-byte aligned rows; a Y that is only 4-byte aligned takes the float-atomic kernel
    const bool pairs_aligned = (d % 2 != 0) || ((uintptr_t)Y % 8 == 0);
    if (atomic_mode && d >= 1 && d <= 4 && pairs_aligned) {
        const Tuning &tu = tuning();
        SampleConst cc = c;
        cc.fast_math = tu.sgd_fast_math;
        switch (tu.sgd_prefetch) {
            case 1: launch_hw<1>(d, grid, Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, cc, N, tu.sgd_plain_far, st); break;
            case 2: launch_hw<2>(d, grid, Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, cc, N, tu.sgd_plain_far, st); break;
            case 8: launch_hw<8>(d, grid, Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, cc, N, tu.sgd_plain_far, st); break;
            default: launch_hw<4>(d, grid, Y, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, cc, N, tu.sgd_plain_far, st); break;
        }
        return;
    }
    switch (d) {
        case 1: k_sgd_epoch<1><<<grid, 256, 0, st>>>(Y, d, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, atomic_mode); break;
        case 2: k_sgd_epoch<2><<<grid, 256, 0, st>>>(Y, d, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, atomic_mode); break;
        case 3: k_sgd_epoch<3><<<grid, 256, 0, st>>>(Y, d, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, atomic_mode); break;
        case 4: k_sgd_epoch<4><<<grid, 256, 0, st>>>(Y, d, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, atomic_mode); break;
        default: k_sgd_epoch<0><<<grid, 256, 0, st>>>(Y, d, row_ptr, col, E, T, chain_row, rng, Qdev, qpart, c, N, atomic_mode); break;
    }
}

void optimize_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t E, float *dY, float *Q,
                  const SgdParams &p, hipStream_t st, float *q_trace, double *epoch_kernel_ms,
                  const SgdWorkspace *ws) {
    if (N < 2 || d < 1) throw HipError("[ncvis_amd] optimize: invalid sizes");
    const int64_t T = p.chains < 1 ? 1 : p.chains;
    DevBuf<int64_t> own_chain_row;
    DevBuf<U128> own_rng;
    DevBuf<float> own_qpart, own_qdev;
    SgdWorkspace w;
    if (ws) {
        w = *ws;
    } else {
        own_chain_row.alloc((size_t)T);
        own_rng.alloc((size_t)T);
        own_qpart.alloc((size_t)T);
        own_qdev.alloc(1 + (size_t)std::max(p.n_epochs, 1));
        w = SgdWorkspace{own_chain_row.p, own_rng.p, own_qpart.p, own_qdev.p};
    }
    k_chain_rows<<<cdiv(T, 256), 256, 0, st>>>(d_row_ptr, N, E, T, w.chain_row, w.rng, p.seed);
    NCVIS_LAUNCH_CHECK();
    NCVIS_HIP_CHECK(hipMemcpyAsync(w.qdev, Q, sizeof(float), hipMemcpyHostToDevice, st));
    dim3 grid((unsigned)cdiv(T, 256));
    // Optional per-epoch kernel timing: one event pair around each epoch kernel on `st`.
    std::vector<hipEvent_t> ev;
    struct EvGuard {
        std::vector<hipEvent_t> &e;
        ~EvGuard() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } evg{ev};
    if (epoch_kernel_ms) {
        ev.resize(2 * (size_t)std::max(p.n_epochs, 0));
        for (auto &x : ev) {
            x = nullptr;
            NCVIS_HIP_CHECK(hipEventCreate(&x));
        }
    }
    for (int epoch = 0; epoch < p.n_epochs; ++epoch) {
        SampleConst c;
        c.a = p.a;
        c.b = p.b;
        c.alpha_Q = p.alpha_Q;
        float fe = (float)epoch / (float)p.n_epochs;
        c.step = p.alpha * (1.f - fe * fe);
        c.noise = (uint32_t)p.n_noise[epoch];
        c.noise_f = (float)p.n_noise[epoch];
        c.fast_math = 0;
        if (epoch_kernel_ms) NCVIS_HIP_CHECK(hipEventRecord(ev[2 * epoch], st));
        launch_epoch(d, grid, dY, d_row_ptr, d_col, E, T, w.chain_row, w.rng, w.qdev, w.qpart, c, (uint64_t)N,
                     p.atomic_updates, st);
        NCVIS_LAUNCH_CHECK();
        if (epoch_kernel_ms) NCVIS_HIP_CHECK(hipEventRecord(ev[2 * epoch + 1], st));
        k_q_reduce<<<1, 1024, 0, st>>>(w.qpart, T, w.qdev, w.qdev + 1, epoch);
        NCVIS_LAUNCH_CHECK();
    }
    NCVIS_HIP_CHECK(hipMemcpyAsync(Q, w.qdev, sizeof(float), hipMemcpyDeviceToHost, st));
    if (q_trace && p.n_epochs > 0)
        NCVIS_HIP_CHECK(
            hipMemcpyAsync(q_trace, w.qdev + 1, sizeof(float) * p.n_epochs, hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (epoch_kernel_ms) {
        double tot = 0.0;
        for (int epoch = 0; epoch < p.n_epochs; ++epoch) {
            float ms = 0.f;
            NCVIS_HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * epoch], ev[2 * epoch + 1]));
            tot += ms;
        }
        *epoch_kernel_ms = tot;
    }
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
