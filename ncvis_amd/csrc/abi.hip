// C ABI of the NCVis HIP library: the handle API behind ncvis::NCVis (src/ncvis.hpp:34-93)
// and the device-pointer stage entry points. Declarations and reference citations are in
// include/ncvis_amd.h.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <atomic>
#include <memory>
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
    int atomic_updates = -1;  // -1: atomics unless a single chain (per-XCD copies by size), 0 plain
                              // stores, 1 device atomics, 2 per-XCD copies of Y
    int64_t chains_override = 0;
    int exact_threads = 0;  // 1: exactly n_threads chains (chain t = the reference's thread t)
    int use_graph = 1;      // replay fits from a captured hipGraph (fit_transform_many: every call;
                            // single fits: from the second call on the same shape and buffers)
    uint64_t gen = 1;       // bumped by every option change
    std::unique_ptr<struct FitBatch> single, batch;  // plans kept for the next fit of the same shapes
    std::string last_error;
    hipStream_t stream = nullptr;
    // stats of the last fit
    double stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t n_edges = 0;
    float final_Q = 0.f;
    int64_t chains_used = 0;
    int sgd_mode = 0;
    int sgd_launches = 0;
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

static ChainRule auto_rule() {
    const Tuning &t = tuning();
    return ChainRule{0, t.edges_per_chain, t.min_chains, t.max_chains, -1};
}

// GPU-sized chain count: E / edges_per_chain, within [min(E / 16, min_chains), max_chains].
// min_chains is off by default: more, shorter chains make a small graph's epochs faster (their
// time is the per-sample latency of a long chain) but move its result away from the reference's
// few-thread Hogwild (the reference's own 1-D two-cluster test fails at E / 16 chains).
static int64_t auto_chains(int64_t E) { return chain_rule(E, auto_rule()); }

// n_threads is the reference's OpenMP thread count. A GPU chain is one such thread, and a few
// sequential chains would leave the GPU idle (one chain at 1M points is ~8e9 dependent samples),
// so the chain count is GPU-sized unless the caller asks for exactly n_threads chains.
static ChainRule rule_of(const ncvis_handle *h) {
    ChainRule r = auto_rule();
    if (h->chains_override > 0)
        r.fixed = h->chains_override;
    else if (h->exact_threads && h->n_threads > 0)
        r.fixed = (int64_t)h->n_threads;
    // atomic_updates 2 asks for the per-XCD copies (a many-chain mode: no lost update across XCDs)
    r.atomic_opt = h->atomic_updates == 2 ? 1 : h->atomic_updates;
    r.replicas = h->atomic_updates == 2 ? 1 : (h->atomic_updates < 0 ? tuning().sgd_replicas : 0);
    return r;
}

__global__ void k_zero_ints(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct FitJob {
    const float *dX;
    size_t N, D;
    float *dY;
    bool operator==(const FitJob &o) const { return dX == o.dX && N == o.N && D == o.D && dY == o.dY; }
};

// ncvis.cpp:364-369 plus this implementation's limits
static void validate_job(const FitJob &j) {
    if (j.N == 0 || j.D == 0)
        throw std::invalid_argument("[ncvis::NCVis::fit_transform] Dataset should have at least one element.");
    if (j.dY == nullptr) throw std::invalid_argument("[ncvis::NCVis::fit_transform] Null pointer provided for output.");
    if (j.dX == nullptr) throw std::invalid_argument("[ncvis_amd] null input pointer");
    if (j.N < 2) throw std::invalid_argument("[ncvis::NCVis::fit_transform] At least two points are required.");
    if (j.N > (size_t)INT32_MAX - 1024) throw std::invalid_argument("[ncvis_amd] more than 2^31 points");
}

// Stage events of an eagerly queued single fit.
struct FitEvents {
    hipEvent_t stage[5];  // start, after knn, after symmetrize, after init, end
    KnnEvents knn;
    std::vector<hipEvent_t> epochs;
    bool per_launch = true;  // record the epoch launches too (not in a graph: ~2,000 event nodes cost
                             // more than the replay saves)
    FitEvents(int n_launches) : epochs(2 * (size_t)std::max(n_launches, 0), nullptr) {
        for (auto &e : stage) e = nullptr;
        knn = KnnEvents{nullptr, nullptr, nullptr};
        for (auto &e : stage) NCVIS_HIP_CHECK(hipEventCreate(&e));
        NCVIS_HIP_CHECK(hipEventCreate(&knn.tiles_begin));
        NCVIS_HIP_CHECK(hipEventCreate(&knn.tiles_end));
        NCVIS_HIP_CHECK(hipEventCreate(&knn.rerank_end));
        for (auto &e : epochs) NCVIS_HIP_CHECK(hipEventCreate(&e));
    }
    ~FitEvents() {
        for (auto e : stage) (void)hipEventDestroy(e);
        (void)hipEventDestroy(knn.tiles_begin);
        (void)hipEventDestroy(knn.tiles_end);
        (void)hipEventDestroy(knn.rerank_end);
        for (auto e : epochs) (void)hipEventDestroy(e);
    }
};

// CSR of a symmetrize group (several data sets' rows concatenated, global indices) split into
// each data set's own row_ptr / col with local indices. One thread per group row.
__global__ void k_split_csr(const FitSeg *__restrict__ s, int B, int64_t base_row, int64_t rows,
                            const int64_t *__restrict__ rp_g, const int32_t *__restrict__ col_g) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t gr = base_row + r;
    int lo = 0, hi = B;  // data set holding global row gr
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (s[m].row_off <= gr)
            lo = m;
        else
            hi = m;
    }
    const FitSeg &g = s[lo];
    const int64_t ro = g.row_off - base_row;  // first group row of the data set
    const int64_t i = r - ro;
    const int64_t eb = rp_g[ro], e0 = rp_g[r], e1 = rp_g[r + 1];
    int64_t *row_ptr = const_cast<int64_t *>(g.row_ptr);
    int32_t *col = const_cast<int32_t *>(g.col);
    row_ptr[i] = e0 - eb;
    if (i == g.N - 1) row_ptr[g.N] = e1 - eb;
    for (int64_t e = e0; e < e1; ++e) col[e - eb] = col_g[e] - (int32_t)ro;
}

// A fit of one or many data sets with one parameter set, every stage queued without host
// synchronisation. Data sets are ordered by their k (n_neighbors clamped to N - 1); each run of
// equal k is one symmetrize group: the kNN tables of the group's data sets are written into one
// table with global row indices and symmetrized in one pass (a row references only rows of its
// own data set, and incoming lists stay in ascending source order, so the result equals each data
// set's own symmetrize), then split into per data set CSR. The chain counts, init and SGD epochs
// of all data sets run segmented, one launch each. The memory is one arena that stays allocated
// for the next fit of the same shapes; a batch can be captured in a hipGraph and replayed while
// its inputs stay the same.
// Multiplier of the affine bijection r -> a r mod N of permuted copies: the integer nearest
// N (sqrt(5) - 1) / 2 that is coprime to N (consecutive rows land about 0.618 N apart, so any run
// of rows spreads evenly over the positions).
static uint64_t perm_multiplier(uint64_t N) {
    uint64_t a = (uint64_t)((double)N * 0.6180339887498949);
    if (a < 1) a = 1;
    auto gcd = [](uint64_t x, uint64_t y) {
        while (y) {
            const uint64_t t = x % y;
            x = y;
            y = t;
        }
        return x;
    };
    while (gcd(a, N) != 1) ++a;
    return a % N;
}

// a^-1 mod N (a coprime to N), extended Euclid
static uint64_t mod_inverse(uint64_t a, uint64_t N) {
    int64_t t = 0, nt = 1, r = (int64_t)N, nr = (int64_t)(a % N);
    while (nr != 0) {
        const int64_t q = r / nr;
        int64_t x = t - q * nt;
        t = nt;
        nt = x;
        x = r - q * nr;
        r = nr;
        nr = x;
    }
    if (t < 0) t += (int64_t)N;
    return (uint64_t)t;
}

struct FitBatch {
    std::vector<FitJob> jobs;   // in the caller's order
    std::vector<int> order;     // descriptor position -> caller's index (sorted by k, stable)
    uint64_t gen = 0;           // handle and tuning generation the plan was made for
    std::vector<int> ks;        // per descriptor position
    std::vector<FitSeg> segs;   // per descriptor position
    struct Group {
        int b0, nb, k;
        int64_t row0, rows;
    };
    std::vector<Group> groups;
    SegTotals tot;
    DevBuf<unsigned char> arena;
    unsigned char *knn_ws = nullptr, *sym_ws = nullptr;
    int32_t *inds = nullptr, *col_g = nullptr;
    float *dists = nullptr, *cdist = nullptr;
    int64_t *rp_g = nullptr;
    int *errs = nullptr;  // [B] kNN flags per descriptor position, then [groups] symmetrize flags
    FitSeg *d_segs = nullptr;
    hipGraphExec_t exec = nullptr;
    std::unique_ptr<FitEvents> ev;  // single fits: the stage and epoch events (recorded by the graph too)
    int eager_runs = 0;             // single fits queued eagerly with the current plan and pointers
    bool no_graph = false;          // a capture failed: single fits stay eager on this plan

    FitBatch() = default;
    FitBatch(const FitBatch &) = delete;
    FitBatch &operator=(const FitBatch &) = delete;
    ~FitBatch() { drop_graph(); }

    bool matches(const std::vector<FitJob> &js, uint64_t g) const { return g == gen && js == jobs; }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        eager_runs = 0;
    }

    void plan(const ncvis_handle *h, const std::vector<FitJob> &js, uint64_t g) {
        drop_graph();
        no_graph = false;
        arena.release();
        jobs = js;
        gen = g;
        const int B = (int)jobs.size();
        const int d = (int)h->d;
        const ChainRule rule = rule_of(h);
        std::vector<int> kk(B);
        for (int i = 0; i < B; ++i) {
            const size_t k0 = (h->n_neighbors < jobs[i].N - 1) ? h->n_neighbors : (jobs[i].N - 1);  // ncvis.cpp:381-382
            kk[i] = (int)(k0 > 0 ? k0 : 1);
        }
        order.resize(B);
        for (int i = 0; i < B; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return kk[x] < kk[y]; });
        ks.assign(B, 0);
        segs.assign(B, FitSeg{});
        groups.clear();
        tot = SegTotals{};
        size_t knn_bytes = 0;
        for (int p = 0; p < B; ++p) {
            const FitJob &j = jobs[order[p]];
            ks[p] = kk[order[p]];
            knn_bytes = std::max(knn_bytes, knn_workspace_bytes((int64_t)j.N, (int)j.D, ks[p], h->dist, h->precision));
            if (groups.empty() || groups.back().k != ks[p]) groups.push_back(Group{p, 0, ks[p], tot.rows, 0});
            groups.back().nb += 1;
            groups.back().rows += (int64_t)j.N;
            segs[p].row_off = tot.rows;
            tot.rows += (int64_t)j.N;
        }
        size_t sym_bytes = 0, gnk = 0, grows = 0;
        for (const Group &gr : groups) {
            sym_bytes = std::max(sym_bytes, symmetrize_workspace_bytes(gr.rows, gr.k));
            gnk = std::max(gnk, (size_t)gr.rows * gr.k);
            if (gr.nb > 1) grows = std::max(grows, (size_t)gr.rows);
        }
        auto a256 = [](size_t x) { return (x + 255) / 256 * 256; };
        size_t o = 0;
        auto take = [&](size_t bytes) {
            const size_t at = o;
            o = a256(o + bytes);
            return at;
        };
        // shared scratch: kNN workspace, the concatenated kNN table of a group, symmetrize
        // workspace and output of a multi-set group
        const size_t o_knn = take(knn_bytes), o_sym = take(sym_bytes), o_inds = take(sizeof(int32_t) * gnk),
                     o_dists = take(sizeof(float) * gnk), o_cdist = take(sizeof(float) * 2 * gnk),
                     o_rpg = take(grows ? sizeof(int64_t) * (grows + 1) : 0),
                     o_colg = take(grows ? sizeof(int32_t) * 2 * gnk : 0);
        // per-XCD copies of Y for the epochs that may run on them (k_seg_dims decides per data set)
        const bool reps = d >= 1 && d <= 4 && rule.replicas != 0 && rule.atomic_opt != 0;
        tot.replicas = reps;
        tot.sub_rows = true;
        struct Off {
            size_t row_ptr, col, y2, part, stats, chain_row, rng, qpart, qdev, yrep, sub_row;
        };
        std::vector<Off> off(B);
        for (int p = 0; p < B; ++p) {
            const FitJob &j = jobs[order[p]];
            FitSeg &v = segs[p];
            const int64_t N = (int64_t)j.N;
            v.N = N;
            v.nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(N, 256)));
            v.T_cap = chain_rule(2 * N * ks[p], rule);  // the rule never decreases with E <= 2 N k
            v.alpha_init = (float)(1. / ks[p]);         // ncvis.cpp:418
            v.pairs_ok = ((uintptr_t)j.dY % 8 == 0) ? 1 : 0;
            v.chain_off = tot.slots;
            tot.slots += v.T_cap;
            tot.max_nb = std::max(tot.max_nb, v.nb);
            off[p].row_ptr = take(sizeof(int64_t) * (N + 1));
            off[p].col = take(sizeof(int32_t) * 2 * N * ks[p]);
            off[p].y2 = take(h->n_init_epochs > 0 ? sizeof(float) * N * d : 0);
            off[p].part = take(sizeof(double) * v.nb * d);
            off[p].stats = take(sizeof(float) * 2 * d);
            off[p].chain_row = take(sizeof(int64_t) * v.T_cap);
            off[p].rng = take(sizeof(U128) * v.T_cap);
            off[p].qpart = take(sizeof(float) * v.T_cap);
            off[p].qdev = take(sizeof(float) * (1 + (size_t)std::max(h->n_epochs, 0)));
            // copies larger than an XCD's L2: rows at permuted positions, negatives by eighths
            const bool perm = reps && tuning().sgd_neg_slices && (size_t)N * d * sizeof(float) > (4u << 20);
            v.perm_a = perm ? perm_multiplier((uint64_t)N) : 0;
            v.perm_ai = perm ? mod_inverse(v.perm_a, (uint64_t)N) : 0;
            tot.permuted = tot.permuted || perm;
            off[p].yrep = take(reps ? sizeof(float) * (kXcds + (perm ? 1 : 0)) * N * d : 0);
            off[p].sub_row = take(sizeof(int64_t) * v.T_cap);
        }
        const size_t o_errs = take(sizeof(int) * (B + groups.size())), o_segs = take(sizeof(FitSeg) * B);
        arena.alloc(o);
        unsigned char *w = arena.p;
        knn_ws = w + o_knn;
        sym_ws = w + o_sym;
        inds = (int32_t *)(w + o_inds);
        dists = (float *)(w + o_dists);
        cdist = (float *)(w + o_cdist);
        rp_g = grows ? (int64_t *)(w + o_rpg) : nullptr;
        col_g = grows ? (int32_t *)(w + o_colg) : nullptr;
        errs = (int *)(w + o_errs);
        d_segs = (FitSeg *)(w + o_segs);
        for (int p = 0; p < B; ++p) {
            FitSeg &v = segs[p];
            v.Y = jobs[order[p]].dY;
            v.Y2 = h->n_init_epochs > 0 ? (float *)(w + off[p].y2) : nullptr;
            v.row_ptr = (const int64_t *)(w + off[p].row_ptr);
            v.col = (const int32_t *)(w + off[p].col);
            v.part = (double *)(w + off[p].part);
            v.stats = (float *)(w + off[p].stats);
            v.chain_row = (int64_t *)(w + off[p].chain_row);
            v.rng = (U128 *)(w + off[p].rng);
            v.qpart = (float *)(w + off[p].qpart);
            v.qdev = (float *)(w + off[p].qdev);
            v.Yrep = reps ? (float *)(w + off[p].yrep) : nullptr;
            v.sub_row = (int64_t *)(w + off[p].sub_row);
        }
    }

    // new input and output pointers for the same shapes (single fits reuse their arena)
    void repoint(const std::vector<FitJob> &js) {
        drop_graph();  // the graph holds the old pointers
        jobs = js;
        for (size_t p = 0; p < segs.size(); ++p) segs[p].Y = jobs[order[p]].dY;
    }

    void upload(hipStream_t st) {
        for (size_t b = 0; b < segs.size(); ++b) seg_set(d_segs + b, segs[b], st);
    }

    // every stage of every data set on `st`; ev (single eager fits only) brackets the stages
    void enqueue(const ncvis_handle *h, hipStream_t st, FitEvents *ev) {
        const int B = (int)jobs.size();
        const int G = (int)groups.size();
        const int d = (int)h->d;
        k_zero_ints<<<cdiv(B + G, 256), 256, 0, st>>>(errs, B + G);
        NCVIS_LAUNCH_CHECK();
        if (ev) record_event(ev->stage[0], st);
        for (int gi = 0; gi < G; ++gi) {
            const Group &gr = groups[gi];
            for (int p = gr.b0; p < gr.b0 + gr.nb; ++p) {
                const FitJob &j = jobs[order[p]];
                const int64_t local = segs[p].row_off - gr.row0;
                knn_enqueue(j.dX, (int64_t)j.N, (int)j.D, gr.k, h->dist, h->precision, inds + local * gr.k,
                            dists + local * gr.k, knn_ws, errs + p, st, ev ? &ev->knn : nullptr, nullptr,
                            (int32_t)local);
            }
            if (ev) record_event(ev->stage[1], st);
            if (gr.nb == 1) {
                symmetrize_enqueue(inds, dists, gr.rows, gr.k, const_cast<int64_t *>(segs[gr.b0].row_ptr),
                                   const_cast<int32_t *>(segs[gr.b0].col), cdist, sym_ws, errs + B + gi, st);
            } else {
                symmetrize_enqueue(inds, dists, gr.rows, gr.k, rp_g, col_g, cdist, sym_ws, errs + B + gi, st);
                k_split_csr<<<cdiv(gr.rows, 256), 256, 0, st>>>(d_segs + gr.b0, gr.nb, gr.row0, gr.rows, rp_g, col_g);
                NCVIS_LAUNCH_CHECK();
            }
        }
        if (ev) record_event(ev->stage[2], st);
        seg_dims(d_segs, B, rule_of(h), d, st);
        seg_init_embedding(d_segs, B, tot, d, h->n_init_epochs, h->random_seed, st);
        if (ev) record_event(ev->stage[3], st);
        SgdParams p;
        p.a = h->a;
        p.b = h->b;
        p.alpha = h->alpha;
        p.alpha_Q = h->alpha_Q;
        p.seed = h->random_seed;
        p.n_epochs = h->n_epochs;
        p.n_noise = h->n_noise.data();
        p.chains = 0;  // per data set, on the device
        p.atomic_updates = h->atomic_updates;
        seg_optimize(d_segs, B, tot, d, 0.f, p, st, ev && ev->per_launch ? ev->epochs.data() : nullptr);  // Q starts at 0 (ncvis.cpp:417)
        if (ev) record_event(ev->stage[4], st);
    }

    // the whole queue of enqueue() as one graph, captured on the handle's own stream (the
    // caller's stream may carry other work); event records become graph nodes
    void capture(const ncvis_handle *h, hipStream_t cap, FitEvents *fev = nullptr) {
        upload(cap);
        NCVIS_HIP_CHECK(hipStreamSynchronize(cap));
        hipGraph_t g = nullptr;
        NCVIS_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
        try {
            if (fev) fev->per_launch = false;
            enqueue(h, cap, fev);
            if (fev) fev->per_launch = true;
        } catch (...) {
            if (fev) fev->per_launch = true;
            (void)hipStreamEndCapture(cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        NCVIS_HIP_CHECK(hipStreamEndCapture(cap, &g));
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        NCVIS_HIP_CHECK(e);
    }

    // error flags of every data set (after the stream has finished); throws for the first
    void check_errors(hipStream_t st, bool name_set) {
        const int B = (int)jobs.size();
        const int G = (int)groups.size();
        std::vector<int> herr(B + G, 0);
        NCVIS_HIP_CHECK(hipMemcpyAsync(herr.data(), errs, sizeof(int) * (B + G), hipMemcpyDeviceToHost, st));
        NCVIS_HIP_CHECK(hipStreamSynchronize(st));
        int first = -1;
        for (int p = 0; p < B; ++p)
            if (herr[p] && (first < 0 || order[p] < order[first])) first = p;
        if (first >= 0) {
            std::string m = "[ncvis_amd] knn: fewer than k+1 finite candidates (non-finite input?)";
            if (name_set) m = "[ncvis_amd] fit_transform_many: data set " + std::to_string(order[first]) + ": " + m;
            throw HipError(m);
        }
        for (int gi = 0; gi < G; ++gi)
            if (herr[B + gi]) throw HipError("[ncvis_amd] symmetrize: neighbour index out of range");
    }
};

// handle parameters and process-wide knobs change a plan: both count generations
static uint64_t plan_gen(const ncvis_handle *h);

struct FitStats {
    double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t n_edges = 0;
    float Q = 0.f;
    int64_t chains = 0;
    int mode = 0;
    int launches = 0;  // epoch kernel launches (epochs x sub-epochs)
};

// Elapsed time between two recorded events; NaN, with the runtime's error cleared, when it has no
// timestamps for them (stage times are statistics, never a reason to fail a fit).
static double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return std::nan("");
    }
    return ms;
}

// One whole fit (ncvis.cpp:354-459) on device buffers, queued on `st`, timed stage by stage.
static void fit_device(ncvis_handle *h, const float *dX, size_t N, size_t D, float *dY, hipStream_t st,
                       FitStats &out);

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
    if (atomic_updates < -1 || atomic_updates > 2) throw std::invalid_argument("[ncvis_amd] atomic_updates: -1, 0, 1 or 2");
    if (atomic_updates == 2 && ws)
        throw std::invalid_argument("[ncvis_amd] atomic_updates 2 (per-XCD copies of Y) allocates; not available here");
    if (atomic_updates == 2 && d > 4) throw std::invalid_argument("[ncvis_amd] atomic_updates 2 needs d <= 4");
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
// pcg64 states, chain start rows, Q copies and sub-epoch rows for up to kNoAllocMaxChains chains,
// the per-epoch Q buffer for up to kNoAllocMaxEpochs epochs, the graph-check flag and the data
// set's descriptor (4.75 MiB per device, held until the process ends). Every call on a device
// shares that device's copy.
constexpr int64_t kNoAllocMaxChains = 1 << 17;
constexpr int kNoAllocMaxEpochs = 1 << 16;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kOffRng = 0;
constexpr size_t kOffChainRow = align_up(kOffRng + sizeof(U128) * kNoAllocMaxChains);
constexpr size_t kOffQpart = align_up(kOffChainRow + sizeof(int64_t) * kNoAllocMaxChains);
constexpr size_t kOffQdev = align_up(kOffQpart + sizeof(float) * kNoAllocMaxChains);
constexpr size_t kOffErr = align_up(kOffQdev + sizeof(float) * (1 + (size_t)kNoAllocMaxEpochs));
constexpr size_t kOffSeg = align_up(kOffErr + sizeof(int));
constexpr size_t kOffSubRow = align_up(kOffSeg + sizeof(FitSeg));
constexpr size_t kScratchBytes = align_up(kOffSubRow + sizeof(int64_t) * kNoAllocMaxChains);

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

static uint64_t plan_gen(const ncvis_handle *h) { return (h->gen << 24) ^ tuning_generation(); }

static void fit_device(ncvis_handle *h, const float *dX, size_t N, size_t D, float *dY, hipStream_t st,
                       FitStats &out) {
    const std::vector<FitJob> js{FitJob{dX, N, D, dY}};
    validate_job(js[0]);
    const uint64_t g = plan_gen(h);
    if (!h->single) h->single.reset(new FitBatch());
    FitBatch &fb = *h->single;
    const bool same_shape = fb.gen == g && fb.jobs.size() == 1 && fb.jobs[0].N == N && fb.jobs[0].D == D &&
                            (((uintptr_t)fb.jobs[0].dY % 8 == 0) == ((uintptr_t)dY % 8 == 0));
    if (!same_shape) {
        fb.plan(h, js, g);
    } else if (!(fb.jobs[0].dX == dX && fb.jobs[0].dY == dY)) {
        fb.repoint(js);  // same arena, new input and output pointers
    }
    const int launches = h->n_epochs * sgd_sub_epochs(fb.tot);
    if (!fb.ev || fb.ev->epochs.size() != 2 * (size_t)launches) {
        fb.drop_graph();  // a captured graph records the old events
        fb.ev.reset(new FitEvents(launches));
    }
    FitEvents &ev = *fb.ev;
    // A fit repeated on the same buffers replays one captured graph: eagerly, its ~2,000 launches
    // (1,000 at 1M: epochs and merges) are issued one host call at a time, and small fits wait on
    // the host. The first fit on a plan and pointers runs eagerly, so one-off fits never pay for a
    // capture.
    if (h->use_graph && fb.eager_runs > 0 && !fb.exec && !fb.no_graph) {
        try {
            fb.capture(h, h->stream, &ev);
        } catch (const std::exception &) {
            // a runtime that cannot capture this queue keeps the fit eager, as with option "graph" 0
            fb.drop_graph();
            fb.no_graph = true;
            fb.eager_runs = 1;
            (void)hipGetLastError();
        }
    }
    const bool replay = h->use_graph && fb.exec;
    if (replay) {
        NCVIS_HIP_CHECK(hipGraphLaunch(fb.exec, st));
    } else {
        fb.upload(st);
        fb.enqueue(h, st, &ev);
        fb.eager_runs += 1;
    }
    fb.check_errors(st, false);  // synchronises st
    FitSeg v;
    NCVIS_HIP_CHECK(hipMemcpy(&v, fb.d_segs, sizeof(FitSeg), hipMemcpyDeviceToHost));
    float Q = 0.f;
    NCVIS_HIP_CHECK(hipMemcpy(&Q, fb.segs[0].qdev, sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; ++i) out.ms[i] = event_ms(ev.stage[i], ev.stage[i + 1]);
    out.ms[4] = out.ms[0] + out.ms[1] + out.ms[2] + out.ms[3];
    out.ms[5] = event_ms(ev.knn.tiles_begin, ev.knn.tiles_end);
    out.ms[6] = event_ms(ev.knn.tiles_end, ev.knn.rerank_end);
    double sgd = replay ? std::nan("") : 0.0;  // a replay records the stages, not every epoch launch
    if (!replay)
        for (int l = 0; l < launches; ++l) sgd += event_ms(ev.epochs[2 * (size_t)l], ev.epochs[2 * (size_t)l + 1]);
    out.ms[7] = sgd;
    out.launches = launches;
    out.n_edges = v.E;
    out.Q = Q;
    out.chains = v.T;
    out.mode = v.mode;
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
        ++h->gen;
        if (k == "knn_precision") {
            int v = (int)value;
            if (v < PREC_FP32 || v > PREC_FP16) throw std::invalid_argument("[ncvis_amd] knn_precision must be 0, 1 or 2");
            h->precision = v;
        } else if (k == "atomic_updates") {
            if (value < -1 || value > 2) throw std::invalid_argument("[ncvis_amd] atomic_updates: -1, 0, 1 or 2");
            h->atomic_updates = (int)value;
        } else if (k == "chains") {
            if (value < 0) throw std::invalid_argument("[ncvis_amd] chains must be >= 0");
            h->chains_override = (int64_t)value;
        } else if (k == "exact_threads") {
            h->exact_threads = value != 0;
        } else if (k == "graph") {
            h->use_graph = value != 0;
        } else {
            throw std::invalid_argument("[ncvis_amd] unknown option '" + k + "'");
        }
    });
}

int ncvis_tune(const char *key, double value) {
    return guarded(g_last_error, [&] {
        std::string k = key ? key : "";
        Tuning &t = tuning();
        bump_tuning_generation();
        if (k == "sgd_prefetch") {
            int v = (int)value;
            if (v != 1 && v != 2 && v != 4 && v != 8) throw std::invalid_argument("[ncvis_amd] sgd_prefetch: 1, 2, 4 or 8");
            t.sgd_prefetch = v;
        } else if (k == "sgd_plain_far") {
            t.sgd_plain_far = value != 0;
        } else if (k == "sgd_fast_math") {
            if (value != 0 && value != 1 && value != 2) throw std::invalid_argument("[ncvis_amd] sgd_fast_math: 0, 1 or 2");
            t.sgd_fast_math = (int)value;
        } else if (k == "sgd_replicas") {
            if (value != 0 && value != 1) throw std::invalid_argument("[ncvis_amd] sgd_replicas: 0 or 1");
            t.sgd_replicas = (int)value;
        } else if (k == "sgd_neg_slices") {
            t.sgd_neg_slices = value != 0;
        } else if (k == "sgd_reread") {
            if (value != -1 && value != 0 && value != 1) throw std::invalid_argument("[ncvis_amd] sgd_reread: -1, 0 or 1");
            t.sgd_reread = (int)value;
        } else if (k == "sgd_edge_flush") {
            t.sgd_edge_flush = value != 0;
        } else if (k == "sgd_sync") {
            if (value < 1 || value > 64) throw std::invalid_argument("[ncvis_amd] sgd_sync: 1..64 sub-epochs");
            t.sgd_sync = (int)value;
        } else if (k == "edges_per_chain") {
            if (value < 1) throw std::invalid_argument("[ncvis_amd] edges_per_chain must be >= 1");
            t.edges_per_chain = (int64_t)value;
        } else if (k == "knn_dma") {
            t.knn_dma = value < 0 ? -1 : (value != 0);
        } else if (k == "knn_split") {
            if (value != 0 && value != 1 && value != 2 && value != 4)
                throw std::invalid_argument("[ncvis_amd] knn_split: 0 (auto), 1, 2 or 4");
            t.knn_split = (int)value;
        } else if (k == "knn_sym") {
            if (value != -1 && value != 0 && !(value >= 1 && value <= 64))
                throw std::invalid_argument("[ncvis_amd] knn_sym: -1 (by size), 0 (list tiles) or a bounding-row "
                                            "stride 1..64");
            t.knn_sym = (int)value;
        } else if (k == "knn_sym_chunk") {
            if (value < 0 || value > 4096) throw std::invalid_argument("[ncvis_amd] knn_sym_chunk: 0..4096 stages");
            t.knn_sym_chunk = (int)value;
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
        h->sgd_mode = fs.mode;
        h->sgd_launches = fs.launches;
    });
}

int ncvis_fit_transform_many(ncvis_handle *h, size_t count, const float *const *dX, const size_t *N, const size_t *D,
                             float *const *dY, int n_streams, void *stream) {
    (void)n_streams;  // the data sets share every segmented launch; no stream split is needed
    if (!h) return NCVIS_EINVAL;
    return guarded(h->last_error, [&] {
        if (count == 0) return;
        if (!dX || !N || !D || !dY) throw std::invalid_argument("[ncvis_amd] fit_transform_many: null array");
        std::vector<FitJob> js(count);
        for (size_t i = 0; i < count; ++i) {
            js[i] = FitJob{dX[i], N[i], D[i], dY[i]};
            try {
                validate_job(js[i]);
            } catch (const std::exception &e) {
                throw std::invalid_argument("[ncvis_amd] fit_transform_many: data set " + std::to_string(i) + ": " +
                                            e.what());
            }
        }
        hipStream_t st = stream ? (hipStream_t)stream : h->stream;
        const uint64_t g = plan_gen(h);
        if (!h->batch) h->batch.reset(new FitBatch());
        FitBatch &fb = *h->batch;
        if (!fb.matches(js, g)) fb.plan(h, js, g);
        if (h->use_graph) {
            if (!fb.exec) fb.capture(h, h->stream);
            NCVIS_HIP_CHECK(hipGraphLaunch(fb.exec, st));
        } else {
            fb.upload(st);
            fb.enqueue(h, st, nullptr);
        }
        fb.check_errors(st, true);
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

int ncvis_get_sgd_mode(const ncvis_handle *h, int *mode, int *launches) {
    if (!h) return NCVIS_EINVAL;
    if (mode) *mode = h->sgd_mode;
    if (launches) *launches = h->sgd_launches;
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
                        (float *)(base + kOffQdev), (FitSeg *)(base + kOffSeg), (int64_t *)(base + kOffSubRow)};
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
