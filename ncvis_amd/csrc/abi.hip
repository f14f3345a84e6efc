// C ABI of the NCVis HIP library: the handle API behind ncvis::NCVis (src/ncvis.hpp:34-93)
// and the device-pointer stage entry points. Declarations and reference citations are in
// include/ncvis_amd.h.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "ncvis_amd.h"

using namespace ncvis_hip;

struct ncvis_handle {
    size_t d = 2, n_threads = 1, n_neighbors = 30, M = 16, ef_construction = 200, random_seed = 42;
    int n_epochs = 50, n_init_epochs = 20;
    float a = 1.f, b = 1.f, alpha = 1.f, alpha_Q = 1.f;
    std::vector<uint64_t> n_noise;
    int dist = DIST_L2;
    int precision = PREC_FP16;
    int atomic_updates = -1;  // -1: atomics unless a single chain
    int64_t chains_override = 0;
    int exact_threads = 0;  // 1: exactly n_threads chains (chain t = the reference's thread t)
    std::string last_error;
    hipStream_t stream = nullptr;
    // stats of the last fit
    double stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t n_edges = 0;
    float final_Q = 0.f;
    int64_t chains_used = 0;
};

static thread_local std::string g_last_error;

namespace ncvis_hip {
Tuning &tuning() {
    static Tuning t;
    return t;
}
}  // namespace ncvis_hip

template <typename F>
static int guarded(std::string &err, F &&f) {
    try {
        f();
        err.clear();
        return NCVIS_OK;
    } catch (const HipError &e) {
        err = e.what();
        return NCVIS_EHIP;
    } catch (const std::invalid_argument &e) {
        err = e.what();
        return NCVIS_EINVAL;
    } catch (const std::exception &e) {
        err = e.what();
        return NCVIS_EINTERNAL;
    } catch (...) {
        err = "[ncvis_amd] unknown error";
        return NCVIS_EINTERNAL;
    }
}

// GPU-sized chain count: E / edges_per_chain, within [min(E / 16, min_chains), max_chains].
// min_chains is off by default: more, shorter chains make a small graph's epochs faster (their
// time is the per-sample latency of a long chain) but move its result away from the reference's
// few-thread Hogwild (the reference's own 1-D two-cluster test fails at E / 16 chains).
static int64_t auto_chains(int64_t E) {
    const Tuning &t = tuning();
    int64_t c = E / std::max<int64_t>(1, t.edges_per_chain);
    c = std::max<int64_t>(c, std::min<int64_t>(E / 16, t.min_chains));
    return std::max<int64_t>(1, std::min<int64_t>(c, t.max_chains));
}

// n_threads is the reference's OpenMP thread count. A GPU chain is one such thread, and a few
// sequential chains would leave the GPU idle (one chain at 1M points is ~8e9 dependent samples),
// so the chain count is GPU-sized unless the caller asks for exactly n_threads chains.
static int64_t resolve_chains(const ncvis_handle *h, int64_t E) {
    if (h->chains_override > 0) return h->chains_override;
    if (h->exact_threads && h->n_threads > 0) return (int64_t)h->n_threads;
    return auto_chains(E);
}

struct FitStats {
    double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t n_edges = 0;
    float Q = 0.f;
    int64_t chains = 0;
};

// One whole fit (ncvis.cpp:354-459) on device buffers, queued on `st`; reads the handle's
// parameters only, so fits of several data sets may run concurrently on their own streams.
static void fit_device(const ncvis_handle *h, const float *dX, size_t N, size_t D, float *dY, hipStream_t st,
                       FitStats &out) {
    // ncvis.cpp:364-369
    if (N == 0 || D == 0)
        throw std::invalid_argument("[ncvis::NCVis::fit_transform] Dataset should have at least one element.");
    if (dY == nullptr)
        throw std::invalid_argument("[ncvis::NCVis::fit_transform] Null pointer provided for output.");
    if (N < 2) throw std::invalid_argument("[ncvis::NCVis::fit_transform] At least two points are required.");
    if (N > (size_t)INT32_MAX - 1024) throw std::invalid_argument("[ncvis_amd] more than 2^31 points");
    // ncvis.cpp:381-382
    size_t kk = (h->n_neighbors < N - 1) ? h->n_neighbors : (N - 1);
    kk = (kk > 0) ? kk : 1;
    const int k = (int)kk;
    const int64_t n = (int64_t)N;
    hipEvent_t ev[5];
    for (auto &e : ev) NCVIS_HIP_CHECK(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 5; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    NCVIS_HIP_CHECK(hipEventRecord(ev[0], st));
    DevBuf<int32_t> inds((size_t)n * k);
    DevBuf<float> dists((size_t)n * k);
    KnnStats ks;
    knn_run(dX, n, (int)D, k, h->dist, h->precision, inds.p, dists.p, st, &ks);
    NCVIS_HIP_CHECK(hipEventRecord(ev[1], st));
    DevBuf<int64_t> row_ptr((size_t)n + 1);
    DevBuf<int32_t> col((size_t)2 * n * k);
    DevBuf<float> cdist((size_t)2 * n * k);
    const int64_t E = symmetrize_run(inds.p, dists.p, n, k, row_ptr.p, col.p, cdist.p, st);
    inds.release();
    dists.release();
    cdist.release();
    NCVIS_HIP_CHECK(hipEventRecord(ev[2], st));
    const int64_t T = resolve_chains(h, E);
    const float init_alpha = (float)(1. / k);  // ncvis.cpp:418
    init_embedding_run(n, (int)h->d, row_ptr.p, col.p, init_alpha, h->n_init_epochs, h->random_seed, T, dY, st);
    NCVIS_HIP_CHECK(hipEventRecord(ev[3], st));
    SgdParams p;
    p.a = h->a;
    p.b = h->b;
    p.alpha = h->alpha;
    p.alpha_Q = h->alpha_Q;
    p.seed = h->random_seed;
    p.n_epochs = h->n_epochs;
    p.n_noise = h->n_noise.data();
    p.chains = T;
    p.atomic_updates = h->atomic_updates >= 0 ? h->atomic_updates : (T > 1 ? 1 : 0);
    float Q = 0.f;  // ncvis.cpp:417
    double epoch_kernel_ms = 0.0;
    optimize_run(n, (int)h->d, row_ptr.p, col.p, E, dY, &Q, p, st, nullptr, &epoch_kernel_ms);
    NCVIS_HIP_CHECK(hipEventRecord(ev[4], st));
    NCVIS_HIP_CHECK(hipEventSynchronize(ev[4]));
    float ms[4];
    for (int i = 0; i < 4; ++i) NCVIS_HIP_CHECK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    for (int i = 0; i < 4; ++i) out.ms[i] = ms[i];
    out.ms[4] = ms[0] + ms[1] + ms[2] + ms[3];
    out.ms[5] = ks.tile_ms;
    out.ms[6] = ks.rerank_ms;
    out.ms[7] = epoch_kernel_ms;
    out.n_edges = E;
    out.Q = Q;
    out.chains = T;
}

// ncvis_optimize_device with its device scratch supplied (ws, err) or allocated (nullptr).
static void optimize_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t n_edges,
                            float *dY, float *Q, int n_epochs, const size_t *n_noise, float a, float b, float alpha,
                            float alpha_Q, uint64_t seed, int64_t n_threads, int atomic_updates, float *q_trace,
                            hipStream_t st, const SgdWorkspace *ws, int *d_err) {
    static_assert(sizeof(size_t) == sizeof(uint64_t), "n_noise is read as uint64_t");
    if (!d_row_ptr || !d_col || !dY || !Q) throw std::invalid_argument("[ncvis_amd] optimize: null pointer");
    if (n_epochs > 0 && !n_noise) throw std::invalid_argument("[ncvis_amd] optimize: null n_noise");
    for (int i = 0; i < n_epochs; ++i)
        if ((uint64_t)n_noise[i] >= 0xFFFFFFFFull) throw std::invalid_argument("[ncvis_amd] n_noise too large");
    if (N < 2) throw std::invalid_argument("[ncvis_amd] optimize: N must be >= 2");
    check_graph(d_row_ptr, (int64_t)N, d_col, n_edges, nullptr, 0, st, d_err);
    SgdParams p;
    p.a = a;
    p.b = b;
    p.alpha = alpha;
    p.alpha_Q = alpha_Q;
    p.seed = seed;
    p.n_epochs = std::max(n_epochs, 0);
    p.n_noise = reinterpret_cast<const uint64_t *>(n_noise);
    p.chains = n_threads < 1 ? auto_chains(n_edges) : n_threads;
    p.atomic_updates = atomic_updates;
    optimize_run((int64_t)N, (int)d, d_row_ptr, d_col, n_edges, dY, Q, p, st, q_trace, nullptr, ws);
}

// ---- scratch of ncvis_optimize_device_noalloc ---------------------------------------------
// One file-scope device array holds the epoch loop's scratch for callers that cannot allocate:
// pcg64 states, chain start rows and Q copies for up to kNoAllocMaxChains chains, the per-epoch
// Q buffer for up to kNoAllocMaxEpochs epochs and the graph-check flag (3.75 MiB per device,
// held until the process ends). Every call on a device shares that device's copy.
constexpr int64_t kNoAllocMaxChains = 1 << 17;
constexpr int kNoAllocMaxEpochs = 1 << 16;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kOffRng = 0;
constexpr size_t kOffChainRow = align_up(kOffRng + sizeof(U128) * kNoAllocMaxChains);
constexpr size_t kOffQpart = align_up(kOffChainRow + sizeof(int64_t) * kNoAllocMaxChains);
constexpr size_t kOffQdev = align_up(kOffQpart + sizeof(float) * kNoAllocMaxChains);
constexpr size_t kOffErr = align_up(kOffQdev + sizeof(float) * (1 + (size_t)kNoAllocMaxEpochs));
constexpr size_t kScratchBytes = align_up(kOffErr + sizeof(int));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the device that owns `st` (the current device for the null
// stream), looked up once per device with that device current. The array's name in host code is
// a host-side stand-in, never a device address: only this pointer is passed to kernels and copies.
static unsigned char *noalloc_scratch(hipStream_t st) {
    static std::mutex mu;
    static std::vector<unsigned char *> ptrs;
    int cur = 0;
    NCVIS_HIP_CHECK(hipGetDevice(&cur));
    int dev = cur;
    if (st) {
        hipDevice_t sd = 0;
        NCVIS_HIP_CHECK(hipStreamGetDevice(st, &sd));
        dev = (int)sd;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (dev >= (int)ptrs.size()) ptrs.resize((size_t)dev + 1, nullptr);
    if (!ptrs[dev]) {
        if (dev != cur) NCVIS_HIP_CHECK(hipSetDevice(dev));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur) NCVIS_HIP_CHECK(hipSetDevice(cur));
        if (e != hipSuccess)
            throw HipError(std::string("[ncvis_amd] cannot locate the no-allocation scratch: ") + hipGetErrorString(e));
        ptrs[dev] = (unsigned char *)p;
    }
    return ptrs[dev];
}

extern "C" {

int ncvis_create(size_t d, size_t n_threads, size_t n_neighbors, size_t M, size_t ef_construction, size_t random_seed,
                 int n_epochs, int n_init_epochs, float a, float b, float alpha, float alpha_Q, const size_t *n_noise,
                 int dist, ncvis_handle **out) {
    return guarded(g_last_error, [&] {
        if (!out) throw std::invalid_argument("[ncvis_amd] ncvis_create: null output handle");
        *out = nullptr;
        if (dist < DIST_L2 || dist > DIST_CORRELATION)
            throw std::invalid_argument("[ncvis::NCVis::buildKNN] Unrecognized distance type.");
        if (d < 1) throw std::invalid_argument("[ncvis_amd] embedding dimension d must be >= 1");
        if (n_epochs < 0 || n_init_epochs < 0) throw std::invalid_argument("[ncvis_amd] epoch counts must be >= 0");
        auto *h = new ncvis_handle();
        h->d = d;
        h->n_threads = n_threads;
        h->n_neighbors = n_neighbors;
        h->M = M;
        h->ef_construction = ef_construction;
        h->random_seed = random_seed;
        h->n_epochs = n_epochs;
        h->n_init_epochs = n_init_epochs;
        h->a = a;
        h->b = b;
        h->alpha = alpha;
        h->alpha_Q = alpha_Q;
        h->dist = dist;
        h->n_noise.resize((size_t)n_epochs);
        for (int i = 0; i < n_epochs; ++i) h->n_noise[i] = n_noise ? (uint64_t)n_noise[i] : 3;  // ncvis.cpp:16-20
        // the epoch kernels count n_noise + 1 samples per edge in 32 bits
        for (uint64_t v : h->n_noise)
            if (v >= 0xFFFFFFFFull) {
                delete h;
                throw std::invalid_argument("[ncvis_amd] n_noise entry too large");
            }
        hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete h;
            throw HipError(std::string("[ncvis_amd] cannot create a HIP stream: ") + hipGetErrorString(e));
        }
        *out = h;
    });
}

void ncvis_destroy(ncvis_handle *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *ncvis_last_error(const ncvis_handle *h) { return h ? h->last_error.c_str() : g_last_error.c_str(); }

const char *ncvis_get_last_error(void) { return g_last_error.c_str(); }

int ncvis_set_option(ncvis_handle *h, const char *key, double value) {
    if (!h) return NCVIS_EINVAL;
    return guarded(h->last_error, [&] {
        std::string k = key ? key : "";
        if (k == "knn_precision") {
            int v = (int)value;
            if (v < PREC_FP32 || v > PREC_FP16) throw std::invalid_argument("[ncvis_amd] knn_precision must be 0, 1 or 2");
            h->precision = v;
        } else if (k == "atomic_updates") {
            h->atomic_updates = (int)value;
        } else if (k == "chains") {
            if (value < 0) throw std::invalid_argument("[ncvis_amd] chains must be >= 0");
            h->chains_override = (int64_t)value;
        } else if (k == "exact_threads") {
            h->exact_threads = value != 0;
        } else {
            throw std::invalid_argument("[ncvis_amd] unknown option '" + k + "'");
        }
    });
}

int ncvis_tune(const char *key, double value) {
    return guarded(g_last_error, [&] {
        std::string k = key ? key : "";
        Tuning &t = tuning();
        if (k == "sgd_prefetch") {
            int v = (int)value;
            if (v != 1 && v != 2 && v != 4 && v != 8) throw std::invalid_argument("[ncvis_amd] sgd_prefetch: 1, 2, 4 or 8");
            t.sgd_prefetch = v;
        } else if (k == "sgd_plain_far") {
            t.sgd_plain_far = value != 0;
        } else if (k == "sgd_fast_math") {
            t.sgd_fast_math = value != 0;
        } else if (k == "edges_per_chain") {
            if (value < 1) throw std::invalid_argument("[ncvis_amd] edges_per_chain must be >= 1");
            t.edges_per_chain = (int64_t)value;
        } else if (k == "knn_dma") {
            t.knn_dma = value < 0 ? -1 : (value != 0);
        } else if (k == "knn_stage_mult") {
            if (value != 1 && value != 2) throw std::invalid_argument("[ncvis_amd] knn_stage_mult: 1 or 2");
            t.knn_stage_mult = (int)value;
        } else if (k == "min_chains") {
            if (value < 1) throw std::invalid_argument("[ncvis_amd] min_chains must be >= 1");
            t.min_chains = (int64_t)value;
        } else if (k == "max_chains") {
            if (value < 1) throw std::invalid_argument("[ncvis_amd] max_chains must be >= 1");
            t.max_chains = (int64_t)value;
        } else {
            throw std::invalid_argument("[ncvis_amd] unknown tuning key '" + k + "'");
        }
    });
}

int ncvis_fit_transform_device(ncvis_handle *h, const float *dX, size_t N, size_t D, float *dY, void *stream) {
    if (!h) return NCVIS_EINVAL;
    return guarded(h->last_error, [&] {
        FitStats fs;
        fit_device(h, dX, N, D, dY, stream ? (hipStream_t)stream : h->stream, fs);
        for (int i = 0; i < 8; ++i) h->stage_ms[i] = fs.ms[i];
        h->n_edges = fs.n_edges;
        h->final_Q = fs.Q;
        h->chains_used = fs.chains;
    });
}

int ncvis_fit_transform_many(ncvis_handle *h, size_t count, const float *const *dX, const size_t *N, const size_t *D,
                             float *const *dY, int n_streams, void *stream) {
    if (!h) return NCVIS_EINVAL;
    return guarded(h->last_error, [&] {
        if (count == 0) return;
        if (!dX || !N || !D || !dY) throw std::invalid_argument("[ncvis_amd] fit_transform_many: null array");
        // inputs may still be in flight on the caller's stream
        NCVIS_HIP_CHECK(hipStreamSynchronize(stream ? (hipStream_t)stream : h->stream));
        const int S = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(n_streams, 1), count));
        std::atomic<size_t> next{0};
        std::mutex mu;
        std::string first_error;
        size_t first_index = count;
        auto worker = [&]() {
            hipStream_t ws = nullptr;
            std::string err;
            int rc = guarded(err, [&] { NCVIS_HIP_CHECK(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking)); });
            if (rc == NCVIS_OK) {
                // stream-ordered allocations: hipFree would wait for every stream of the device
                alloc_stream() = ws;
                for (size_t i = next++; i < count; i = next++) {
                    FitStats fs;
                    rc = guarded(err, [&] { fit_device(h, dX[i], N[i], D[i], dY[i], ws, fs); });
                    if (rc != NCVIS_OK) {
                        std::lock_guard<std::mutex> lk(mu);
                        if (i < first_index) {
                            first_index = i;
                            first_error = err;
                        }
                    }
                }
                (void)hipStreamSynchronize(ws);
                alloc_stream() = nullptr;
                (void)hipStreamDestroy(ws);
            } else {
                std::lock_guard<std::mutex> lk(mu);
                if (first_error.empty()) first_error = err;
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < S; ++t) pool.emplace_back(worker);
        worker();
        for (auto &t : pool) t.join();
        if (!first_error.empty()) {
            if (first_index < count)
                throw std::runtime_error("[ncvis_amd] fit_transform_many: data set " + std::to_string(first_index) +
                                         ": " + first_error);
            throw std::runtime_error(first_error);
        }
    });
}

int ncvis_fit_transform(ncvis_handle *h, const float *X, size_t N, size_t D, float *Y) {
    if (!h) return NCVIS_EINVAL;
    int rc = guarded(h->last_error, [&] {
        if (N == 0 || D == 0)
            throw std::invalid_argument("[ncvis::NCVis::fit_transform] Dataset should have at least one element.");
        if (Y == nullptr)
            throw std::invalid_argument("[ncvis::NCVis::fit_transform] Null pointer provided for output.");
        if (X == nullptr) throw std::invalid_argument("[ncvis_amd] null input pointer");
    });
    if (rc) return rc;
    DevBuf<float> dX, dY;
    rc = guarded(h->last_error, [&] {
        dX.alloc(N * D);
        dY.alloc(N * h->d);
        NCVIS_HIP_CHECK(hipMemcpyAsync(dX.p, X, sizeof(float) * N * D, hipMemcpyHostToDevice, h->stream));
    });
    if (rc) return rc;
    rc = ncvis_fit_transform_device(h, dX.p, N, D, dY.p, h->stream);
    if (rc) return rc;
    return guarded(h->last_error, [&] {
        NCVIS_HIP_CHECK(hipMemcpyAsync(Y, dY.p, sizeof(float) * N * h->d, hipMemcpyDeviceToHost, h->stream));
        NCVIS_HIP_CHECK(hipStreamSynchronize(h->stream));
    });
}

int ncvis_get_stats(const ncvis_handle *h, double *stage_ms8, int64_t *n_edges, float *final_Q, int64_t *chains) {
    if (!h) return NCVIS_EINVAL;
    if (stage_ms8)
        for (int i = 0; i < 8; ++i) stage_ms8[i] = h->stage_ms[i];
    if (n_edges) *n_edges = h->n_edges;
    if (final_Q) *final_Q = h->final_Q;
    if (chains) *chains = h->chains_used;
    return NCVIS_OK;
}

int ncvis_knn_device(const float *dX, size_t N, size_t D, size_t k, int dist, int precision, int32_t *d_inds,
                     float *d_dists, void *stream) {
    return guarded(g_last_error, [&] {
        if (!dX || !d_inds || !d_dists) throw std::invalid_argument("[ncvis_amd] knn: null pointer");
        if (dist < DIST_L2 || dist > DIST_CORRELATION) throw std::invalid_argument("[ncvis_amd] knn: bad distance");
        if (N < 2 || D < 1 || k < 1 || k > N - 1) throw std::invalid_argument("[ncvis_amd] knn: need 1 <= k <= N-1");
        knn_run(dX, (int64_t)N, (int)D, (int)k, dist, precision, d_inds, d_dists, (hipStream_t)stream);
    });
}

int ncvis_symmetrize_device(const int32_t *d_inds, const float *d_dists, size_t N, size_t k, int64_t *d_row_ptr,
                            int32_t *d_col, float *d_dist, int64_t *n_edges, void *stream) {
    return guarded(g_last_error, [&] {
        if (!d_inds || !d_dists || !d_row_ptr || !d_col || !d_dist)
            throw std::invalid_argument("[ncvis_amd] symmetrize: null pointer");
        if (N < 1 || k < 1) throw std::invalid_argument("[ncvis_amd] symmetrize: empty table");
        int64_t E = symmetrize_run(d_inds, d_dists, (int64_t)N, (int)k, d_row_ptr, d_col, d_dist, (hipStream_t)stream);
        if (n_edges) *n_edges = E;
    });
}

int ncvis_init_embedding_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, float alpha,
                                int n_init_epochs, uint64_t seed, int64_t n_threads, float *dY, void *stream) {
    return guarded(g_last_error, [&] {
        if (!d_row_ptr || !dY) throw std::invalid_argument("[ncvis_amd] init: null pointer");
        if (N < 1) throw std::invalid_argument("[ncvis_amd] init: N must be >= 1");
        int64_t E = 0;
        NCVIS_HIP_CHECK(hipMemcpyAsync(&E, d_row_ptr + N, sizeof(int64_t), hipMemcpyDeviceToHost,
                                       (hipStream_t)stream));
        NCVIS_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
        check_graph(d_row_ptr, (int64_t)N, d_col, E, nullptr, 0, (hipStream_t)stream);
        init_embedding_run((int64_t)N, (int)d, d_row_ptr, d_col, alpha, n_init_epochs, seed,
                           n_threads < 1 ? 1 : n_threads, dY, (hipStream_t)stream);
    });
}

int ncvis_optimize_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t n_edges,
                          float *dY, float *Q, int n_epochs, const size_t *n_noise, float a, float b, float alpha,
                          float alpha_Q, uint64_t seed, int64_t n_threads, int atomic_updates, float *q_trace,
                          void *stream) {
    return guarded(g_last_error, [&] {
        optimize_device(N, d, d_row_ptr, d_col, n_edges, dY, Q, n_epochs, n_noise, a, b, alpha, alpha_Q, seed,
                        n_threads, atomic_updates, q_trace, (hipStream_t)stream, nullptr, nullptr);
    });
}

int ncvis_optimize_device_noalloc(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col,
                                  int64_t n_edges, float *dY, float *Q, int n_epochs, const size_t *n_noise, float a,
                                  float b, float alpha, float alpha_Q, uint64_t seed, int64_t n_threads,
                                  int atomic_updates, float *q_trace, void *stream) {
    return guarded(g_last_error, [&] {
        const int64_t T = n_threads < 1 ? auto_chains(n_edges) : n_threads;
        if (T > kNoAllocMaxChains)
            throw std::invalid_argument("[ncvis_amd] optimize_noalloc: " + std::to_string(T) +
                                        " chains exceed the scratch limit of " + std::to_string(kNoAllocMaxChains));
        if (n_epochs > kNoAllocMaxEpochs)
            throw std::invalid_argument("[ncvis_amd] optimize_noalloc: " + std::to_string(n_epochs) +
                                        " epochs exceed the scratch limit of " + std::to_string(kNoAllocMaxEpochs));
        unsigned char *base = noalloc_scratch((hipStream_t)stream);
        SgdWorkspace ws{(int64_t *)(base + kOffChainRow), (U128 *)(base + kOffRng), (float *)(base + kOffQpart),
                        (float *)(base + kOffQdev)};
        optimize_device(N, d, d_row_ptr, d_col, n_edges, dY, Q, n_epochs, n_noise, a, b, alpha, alpha_Q, seed, T,
                        atomic_updates, q_trace, (hipStream_t)stream, &ws, (int *)(base + kOffErr));
    });
}

int ncvis_nce_loss_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t n_edges,
                          const int32_t *d_neg, int n, const float *dY, float Q, float a, float b, double *d_loss,
                          float *d_gradY, double *d_gradQ, void *stream) {
    return guarded(g_last_error, [&] {
        if (!d_row_ptr || !d_col || !d_neg || !dY || !d_loss)
            throw std::invalid_argument("[ncvis_amd] nce_loss: null pointer");
        if (N < 1 || n < 1) throw std::invalid_argument("[ncvis_amd] nce_loss: needs N >= 1 and n >= 1");
        check_graph(d_row_ptr, (int64_t)N, d_col, n_edges, d_neg, n_edges * (int64_t)n, (hipStream_t)stream);
        nce_loss_run((int64_t)N, (int)d, d_row_ptr, d_col, n_edges, d_neg, n, dY, Q, a, b, d_loss, d_gradY, d_gradQ,
                     (hipStream_t)stream);
    });
}

}  // extern "C"
