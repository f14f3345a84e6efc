// Internal host-side launchers of the NCVis HIP kernels (not part of the C ABI).
#pragma once

#include "common.h"

namespace ncvis_hip {

// error flags a queued fit leaves in its device error word (read once at the end)
constexpr int kErrKnnCandidates = 1;  // a row had fewer than k+1 finite candidates
constexpr int kErrSymRange = 2;       // a kNN index outside [0, N)

// Process-wide tuning knobs (ncvis_tune in the C ABI). Defaults are the measured best on MI355X;
// none of them changes the result of a single-chain (n_threads = 1) fit.
struct Tuning {
    int sgd_prefetch = 1;            // rows fetched together per chain in the many-chain epoch (1,2,4,8)
    int sgd_plain_far = 0;           // 1: far-endpoint updates with plain stores instead of atomics
    int sgd_replicas = 1;            // many-chain epochs on per-XCD copies of Y (mode 3): 1 when the
                                     // atomics are automatic, 0 never (device atomics, mode 2)
    int sgd_sync = 16;               // sub-epochs per epoch: every chain writes its source row out and,
                                     // with copies, the copies are merged after each
    int sgd_neg_slices = 1;          // per-XCD copies beyond an XCD L2: rows at permuted positions, negatives by eighths
    int sgd_reread = -1;             // per-XCD copies: read the far row again right before its store
                                     // (1 always, 0 never, -1 unless the copies are permuted)
    int sgd_edge_flush = 0;          // per-XCD copies: write the source row's update after every edge
                                     // (0: at row changes and sub-epoch ends)
    int sgd_fast_math = 2;           // many-chain epochs: 2 hardware log2/exp2/reciprocal, 1 single-
                                     // precision powf/expf, 0 double rounded once (as the exact one-chain
                                     // mode always does)
    int64_t edges_per_chain = 32;    // auto chain count = E / edges_per_chain,
    int64_t min_chains = 1;          // ... at least min(E / 16, min_chains) (1: off) ...
    int64_t max_chains = 1 << 17;    // ... and at most this
    int knn_dma = -1;                // FP32 kNN tiles: stages by LDS DMA (1 or auto -1) or registers (0)
    int knn_stage_mult = 2;          // 16-bit kNN (k + 1 <= 16): LDS stage rows x1 or x2
    int knn_split = 0;               // candidate split of untuned searches (0: by workgroup count)
    int knn_sym = -1;                // 16-bit kNN, k + 1 <= 16: symmetric threshold search with row
                                     // bounds from every knn_sym-th row (0: list tiles; -1: by size)
    int knn_sym_chunk = 128;         // symmetric search: candidate stages per chunk of the chunk-major
                                     // work order (0: each workgroup sweeps two query blocks' ranges)
};
Tuning &tuning();

// ---- kNN (knn.hip) ---------------------------------------------------------------------
// Result-invariant launch choices of the candidate tiles: the candidate split (grid.y ranges of
// the candidate sweep, each keeping lists of its own, so small searches fill the GPU; every
// range's lists go to the exact rerank), the 16-bit LDS stage multiplier and FP32 LDS DMA.
struct TileVariant {
    int split = 1;
    int stage_mult = 2;
    int dma = 1;
    int sym = 0;  // > 0: symmetric threshold search (knn_sym.hip), row bounds from every sym-th row
};

// Workspace of the symmetric threshold search (knn_sym.hip).
struct SymBuffers {
    float *T;       // [Npad] score bound of each row
    int *cnt;       // [Npad + 1] bucket fill counts, then the fallback row count
    uint2 *bucket;  // [N, sym_bucket_entries()] (score bits, label)
    int *ovf;       // [N] rows left to the fallback
};
size_t sym_bucket_entries();
// Candidate stage of the symmetric search with row bounds from every stride-th row: writes cand
// [N, 32] (ids, -1 padded) for the rerank.
void knn_sym_enqueue(int prec, int DK, const void *Xh, const float *cn, int64_t N, int64_t Npad, float beta,
                     int K1, int stride, int split, const SymBuffers &w, int32_t *cand, hipStream_t st);

// Exact k-nearest-neighbour table in the reference's findKNN row format: for each row the k+1
// smallest (distance, label) pairs are found, the smallest is dropped and the remaining k are
// written farthest first. `prec` selects the candidate GEMM precision; the final distances are
// always exact FP32 (squared L2, or 1 - <x,y> for the inner-product family).
struct KnnStats {
    double tile_ms = 0;    // candidate GEMM + top-k kernel time
    double rerank_ms = 0;  // exact rerank time
    int used_mfma = 0;     // 1 if the MFMA tile kernel ran, 0 if the generic kernel ran
};
void knn_run(const float *dX, int64_t N, int D, int k, int dist, int prec, int32_t *d_inds, float *d_dists,
             hipStream_t st, KnnStats *stats = nullptr);
// Events recorded around the candidate tiles and the rerank (optional).
struct KnnEvents {
    hipEvent_t tiles_begin, tiles_end, rerank_end;
};
// The same search queued on `st` with no host synchronisation and no allocation: ws holds
// knn_workspace_bytes(N, D, k, dist, prec); a row left with fewer than k+1 finite candidates sets
// kErrKnnCandidates in *d_err (which the caller zeroes). label_off is added to every written
// neighbour index (tables of several data sets concatenated into one).
size_t knn_workspace_bytes(int64_t N, int D, int k, int dist, int prec);
void knn_enqueue(const float *dX, int64_t N, int D, int k, int dist, int prec, int32_t *d_inds, float *d_dists,
                 void *ws, int *d_err, hipStream_t st, const KnnEvents *ev = nullptr,
                 const TileVariant *force = nullptr, int32_t label_off = 0);
// Workspace of a search with a forced variant (the autotuner's measurements).
size_t knn_workspace_bytes_for(int64_t N, int D, int k, int dist, int prec, const TileVariant &tv);
// The variant a forced one becomes once the search's limits clamp it.
TileVariant knn_effective_variant(int64_t N, int D, int k, int dist, int prec, const TileVariant &tv);

// Process-wide generation of everything a fit plan depends on besides its handle (tuning knobs,
// the autotuning table); bumped by every change.
uint64_t tuning_generation();
void bump_tuning_generation();

// The variant the autotuning table holds for this search (false: none).
bool tuned_variant(int prec, int DK, int KL, bool wide, int64_t N, TileVariant &out);
// The variant knn_enqueue uses for this search.
TileVariant knn_variant(int64_t N, int D, int k, int dist, int prec);

// Candidate-tile launches on the packed matrix Xh [Npad, 32*DK] and offsets cn [Npad], writing,
// for each of the `split` candidate ranges, the lane lists of KL ids per query to
// cand [N, split, lists * KL] (one translation unit each).
void knn_tiles16(int prec, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                 int32_t *cand, hipStream_t st, const TileVariant &tv);  // KL = 16 (knn_t16.hip)
// KL = 32 or 64, one translation unit per GEMM precision / FP32 staging (knn_q1_*.hip)
void knn_tiles_q1_f32dma(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                         int32_t *cand, hipStream_t st, int split);
void knn_tiles_q1_f32(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                      int32_t *cand, hipStream_t st, int split);
void knn_tiles_q1_bf16(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                       int32_t *cand, hipStream_t st, int split);
void knn_tiles_q1_fp16(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                       int32_t *cand, hipStream_t st, int split);
void knn_tiles_wide(int prec, int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn,
                    float beta, int32_t *cand, hipStream_t st, int split);  // 16-bit, DK 10/12/16 (knn_wide.hip)
// FP32 rows of 257..512 elements (DK 10/12/16, KL 16/32): 16x16x4 tiles, four lists of KL per
// query and range, Npad a multiple of 64 (knn_wide_f32.hip)
void knn_tiles_wide_f32(int KL, int DK, int64_t N, int64_t Npad, const void *Xh, const float *cn, float beta,
                        int32_t *cand, hipStream_t st, int split);

// ---- symmetrize (symmetrize.hip) ---------------------------------------------------------
// KNNTable::symmetrize at one thread, bit-exact, then flattened to CSR (= build_edges order).
// d_col / d_dist must hold 2*N*k entries; d_row_ptr N+1. Synchronous; returns E.
int64_t symmetrize_run(const int32_t *d_inds, const float *d_dists, int64_t N, int k, int64_t *d_row_ptr,
                       int32_t *d_col, float *d_dist, hipStream_t st);
// The same queued on `st` with no host synchronisation and no allocation: ws holds
// symmetrize_workspace_bytes(N, k); E ends up in d_row_ptr[N]; an index outside [0, N) sets
// kErrSymRange in *d_err (which the caller zeroes) and is left out of the incoming lists.
size_t symmetrize_workspace_bytes(int64_t N, int k);
void symmetrize_enqueue(const int32_t *d_inds, const float *d_dists, int64_t N, int k, int64_t *d_row_ptr,
                        int32_t *d_col, float *d_dist, void *ws, int *d_err, hipStream_t st);

// ---- embedding (embed.hip) ---------------------------------------------------------------
struct SgdParams {
    float a, b, alpha, alpha_Q;
    uint64_t seed;
    int n_epochs;
    const uint64_t *n_noise;  // host array [n_epochs]
    int64_t chains;           // T: number of sequential SGD chains (the reference's n_threads)
    int atomic_updates;       // 1: Y updates with atomics (Hogwild without lost updates)
};

// Chain count of a data set with E edges (the reference's thread count on the GPU): `fixed` if
// positive, else E / edges_per_chain within [min(E / 16, min_chains), max_chains], at least 1.
// atomic_opt: -1 atomics exactly when more than one chain runs, 0 never, 1 always.
struct ChainRule {
    int64_t fixed;
    int64_t edges_per_chain, min_chains, max_chains;
    int atomic_opt;
    int replicas = 0;  // 1: many-chain epochs on per-XCD copies of Y (when d <= 4), 0: device atomics
};
__host__ __device__ inline int64_t chain_rule(int64_t E, const ChainRule &r) {
    if (r.fixed > 0) return r.fixed;
    int64_t c = E / (r.edges_per_chain > 0 ? r.edges_per_chain : 1);
    const int64_t lo = (E / 16 < r.min_chains) ? E / 16 : r.min_chains;
    if (c < lo) c = lo;
    if (c > r.max_chains) c = r.max_chains;
    return c < 1 ? 1 : c;
}

// One data set of a fit, in device memory. The init and epoch kernels walk the concatenation of
// B data sets (rows at row_off, chain slots at chain_off, both ascending).
struct FitSeg {
    float *Y;                  // [N, d] embedding
    float *Y2;                 // [N, d] init double buffer (null without init epochs)
    const int64_t *row_ptr;    // [N + 1]
    const int32_t *col;        // [E]
    int64_t N;
    int64_t row_off;           // first row of this data set in the concatenated rows
    int64_t chain_off;         // first chain slot of this data set
    int64_t T_cap;             // chain slots reserved (T never exceeds it)
    int64_t *chain_row;        // [T_cap] row of each chain's first edge
    U128 *rng;                 // [T_cap] pcg64 state of each chain
    float *qpart;              // [T_cap] each chain's Q copy
    float *qdev;               // [1 + n_epochs]: Q, then Q after each epoch
    double *part;              // [nb * d] init statistics partial sums
    float *stats;              // [2 d] init mean and sigma
    float alpha_init;          // propagation weight 1 / k (ncvis.cpp:418)
    int nb;                    // blocks of the init statistics (min(1024, ceil(N / 256)))
    int pairs_ok;              // Y is 8-byte aligned (64-bit compare-and-swap on pairs)
    int64_t E, T;              // edges and chains (k_seg_dims, or the host for stage calls)
    int mode;                  // 0 exact plain stores, 1 exact with float atomics, 2 many-chain with
                               // device atomics, 3 many-chain on per-XCD copies of Y
    float *Yrep;               // [kXcds, N, d] per-XCD copies of Y (mode 3), or null
    int64_t *sub_row;          // [T_cap] row of each chain's last edge between sub-epochs (null: one
                               // launch per epoch)
    uint64_t perm_a, perm_ai;  // mode 3, copies larger than an XCD's L2 (sgd_neg_slices): the copies
                               // hold row r at position perm_a r mod N (perm_ai inverts it) and
                               // Yrep holds kXcds + 1 arrays (the last: the merges' base); 0: rows
                               // in place
};

// XCDs (L2 domains) of the part; the per-XCD copies of Y are indexed by the hardware XCC id.
constexpr int kXcds = 8;

struct SegTotals {
    int64_t rows = 0;       // sum of N
    int64_t slots = 0;      // sum of T_cap
    int max_nb = 1;         // largest nb
    bool replicas = false;  // some data set has per-XCD copies of Y (merged after every sub-epoch)
    bool sub_rows = false;  // every data set has sub_row: epochs may run as several sub-epochs
    bool permuted = false;  // some data set has permuted copies (perm_a != 0)
};

// Writes one descriptor from the host value (a kernel argument: no copy, capturable).
void seg_set(FitSeg *d_seg, const FitSeg &v, hipStream_t st);
// E, T and mode of every data set from its graph on the device.
void seg_dims(FitSeg *d_segs, int B, const ChainRule &rule, int d, hipStream_t st);
// Random layout + n_init_epochs propagation epochs of every data set (T must be set).
void seg_init_embedding(const FitSeg *d_segs, int B, const SegTotals &tot, int d, int n_init_epochs, uint64_t seed,
                        hipStream_t st);
// All SGD epochs of every data set from Q0 (E, T, mode must be set); no host synchronisation.
// Each epoch runs as sgd_sub_epochs(tot) launches of the epoch kernel (plus the copies' merges);
// epoch_events (optional, 2 * n_epochs * sgd_sub_epochs(tot) events) bracket each launch.
int sgd_sub_epochs(const SegTotals &tot);
void seg_optimize(const FitSeg *d_segs, int B, const SegTotals &tot, int d, float Q0, const SgdParams &p,
                  hipStream_t st, hipEvent_t *epoch_events);
// summed time of the epoch kernel launches between event pairs
double epoch_time_ms(const hipEvent_t *ev, int n_launches);

// One data set, synchronous (the stage entry points): random layout + propagation epochs with
// `chains` threads' static schedule of the random draws.
void init_embedding_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, float alpha,
                        int n_init_epochs, uint64_t seed, int64_t chains, float *dY, hipStream_t st);

// Device scratch of the epoch loop, for callers that bring their own memory: chain_row [T],
// rng [T] (16-byte aligned), qpart [T], qdev [1 + n_epochs], seg [1], sub_row [T].
struct SgdWorkspace {
    int64_t *chain_row;
    U128 *rng;
    float *qpart;
    float *qdev;
    FitSeg *seg;
    int64_t *sub_row;  // [T]
};

// One data set with host-known E and chain count, synchronous; Q (host in/out). `q_trace`
// (optional, host [n_epochs]) receives Q per epoch; `epoch_kernel_ms` (optional, host) the summed
// event time of the epoch kernels alone. `ws` (optional) supplies the device scratch; without it
// the scratch is allocated here.
void optimize_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t E, float *dY,
                  float *Q, const SgdParams &p, hipStream_t st, float *q_trace = nullptr,
                  double *epoch_kernel_ms = nullptr, const SgdWorkspace *ws = nullptr);

// Validates a caller-supplied graph before a kernel indexes with it: row_ptr [N+1] starts at 0,
// ends at E and never decreases; col [E] and neg [n_neg] (may be null when n_neg = 0) lie in
// [0, N). Throws std::invalid_argument otherwise. Synchronises `st`. `d_err` (optional) is one
// device int of caller memory for the result flags; without it one is allocated here.
void check_graph(const int64_t *d_row_ptr, int64_t N, const int32_t *d_col, int64_t E, const int32_t *d_neg,
                 int64_t n_neg, hipStream_t st, int *d_err = nullptr);

// NCE loss and adjoint over an explicit negative table neg[E, n] (int32). dY (float, [N,d],
// accumulated) and dQ / loss (double) are device pointers; dY may be null for loss only.
void nce_loss_run(int64_t N, int d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t E,
                  const int32_t *d_neg, int n, const float *dY_in, float Q, float a, float b, double *d_loss,
                  float *d_gradY, double *d_gradQ, hipStream_t st);

}  // namespace ncvis_hip
