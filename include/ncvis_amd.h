/*
 * C ABI of the MI355X-native NCVis fit path (libncvis_amd.so, gfx950).
 *
 * Plain pointers and sizes only. Streams are passed as `void*` holding a hipStream_t (NULL =
 * the library's own stream for handle calls, the default stream for stage calls). Every entry
 * point returns 0 on success and a negative status on failure; the message is available from
 * ncvis_last_error() (handle calls) or ncvis_get_last_error() (stage calls, per thread).
 *
 * Reference interfaces these replace (paths relative to the reference repository):
 *   ncvis_create / ncvis_destroy  -> ncvis::NCVis::NCVis / ~NCVis (src/ncvis.hpp:54-55,
 *                                    src/ncvis.cpp:12-30)
 *   ncvis_fit_transform           -> ncvis::NCVis::fit_transform (src/ncvis.hpp:66,
 *                                    src/ncvis.cpp:354-459), host buffers
 *   ncvis_knn_device              -> buildKNN + findKNN (src/ncvis.cpp:72-153)
 *   ncvis_symmetrize_device       -> KNNTable::symmetrize + build_edges (src/knntable.cpp:24-99,
 *                                    src/ncvis.cpp:155-173)
 *   ncvis_init_embedding_device   -> init_embedding (src/ncvis.cpp:184-273)
 *   ncvis_optimize_device         -> optimize (src/ncvis.cpp:275-352)
 *   ncvis_nce_loss_device         -> no reference function: the NCE objective whose gradient
 *                                    src/ncvis.cpp:304-338 applies in place
 */
#ifndef NCVIS_AMD_H
#define NCVIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distance codes: the ordinals of ncvis::Distance (src/ncvis.hpp:27-32). */
enum ncvis_distance { NCVIS_SQUARED_L2 = 0, NCVIS_INNER_PRODUCT = 1, NCVIS_COSINE = 2, NCVIS_CORRELATION = 3 };

/* Candidate-GEMM precision of the kNN search; final distances are always exact FP32. */
enum ncvis_precision { NCVIS_PREC_FP32 = 0, NCVIS_PREC_BF16 = 1, NCVIS_PREC_FP16 = 2 };

/* Status codes */
enum ncvis_status { NCVIS_OK = 0, NCVIS_EINVAL = -1, NCVIS_EHIP = -2, NCVIS_EINTERNAL = -3 };

typedef struct ncvis_handle ncvis_handle;

/* ncvis::NCVis::NCVis (src/ncvis.hpp:54). n_noise may be NULL (3 per epoch, src/ncvis.cpp:16-20)
 * and is copied; entries must be below 2^32 - 1. n_threads is the reference's OpenMP thread
 * count. Each GPU SGD chain plays one such thread (its own pcg64(random_seed + chain), a static
 * chunk of the edges, a private Q copy). By default the chain count is GPU-sized (E / 32, at most
 * 2^17) whatever n_threads says, because a few sequential chains leave the GPU idle. Option
 * "exact_threads" = 1 runs exactly n_threads chains instead; n_threads = 1 then reproduces the
 * reference at one thread. M and ef_construction are accepted for signature compatibility; the
 * search is exact. */
int ncvis_create(size_t d, size_t n_threads, size_t n_neighbors, size_t M, size_t ef_construction,
                 size_t random_seed, int n_epochs, int n_init_epochs, float a, float b, float alpha,
                 float alpha_Q, const size_t *n_noise, int dist, ncvis_handle **out);

void ncvis_destroy(ncvis_handle *h);

/* Message of the last failed call on this handle ("" if none). Valid until the next call. */
const char *ncvis_last_error(const ncvis_handle *h);

/* Options (no reference counterpart):
 *   "knn_precision"   NCVIS_PREC_*  (default NCVIS_PREC_FP16)
 *   "atomic_updates"  -1 (default): one chain with plain stores; many chains on per-XCD copies
 *                     of Y (d <= 4 and tuning "sgd_replicas" 1, else device atomics);
 *                     0: plain read-modify-write stores like the reference; 1: device-scope
 *                     atomics on Y (no update lost); 2: per-XCD copies of Y (d <= 4): each XCD
 *                     updates its own copy with plain stores in its L2 (a chain of the same XCD
 *                     can overwrite a concurrent update of a row, the reference's Hogwild race),
 *                     and the copies are merged after every one of "sgd_sync" sub-epochs
 *   "exact_threads"   1: run exactly n_threads chains (default 0: GPU-sized chain count)
 *   "chains"          override the SGD chain count (0 = none)
 *   "graph"           1 (default): ncvis_fit_transform_many captures its work in a hipGraph and
 *                     replays it; a device-pointer fit repeated on the same buffers and shape is
 *                     captured at its second call and replayed from then on (stage times included);
 *                     0: every fit is queued eagerly */
int ncvis_set_option(ncvis_handle *h, const char *key, double value);

/* Process-wide tuning knobs (no reference counterpart). None changes a single-chain fit; the kNN
 * knobs never change the kNN table. Keys: "sgd_prefetch" (1,2,4,8), "sgd_plain_far" (0/1),
 * "sgd_fast_math" (many-chain sample math: 2 hardware log2/exp2/reciprocal (default), 1 powf/expf,
 * 0 double), "sgd_replicas" (1: many-chain epochs on per-XCD copies of Y when the
 * atomics are automatic, 0: device atomics), "sgd_sync" (sub-epochs per epoch, 1..64: every chain
 * writes its source row out after each, and the copies are merged), "sgd_edge_flush" (0/1: with
 * copies, the source row is written after every edge), "sgd_reread" (per-XCD copies: 1 read the
 * far row again right before its store, 0 store the prefetched row minus the update, -1 (default)
 * 0 on permuted copies, 1 otherwise), "sgd_neg_slices" (1 (default): copies larger than an XCD's
 * L2 hold the rows at permuted positions and a launch draws an XCD's negatives from one eighth of
 * them; 0: rows in place, negatives over all rows), "edges_per_chain" (auto chain count = E / this, at least
 * min(E / 16, "min_chains"), at most "max_chains"),
 * "knn_dma" (FP32 tiles: -1 auto = 1 LDS DMA, 0 register-staged), "knn_stage_mult" (16-bit
 * tiles: 1 or 2 x the LDS stage rows), "knn_split" (candidate ranges per query block: 0 auto,
 * else 1, 2 or 4; each range keeps lists of its own, so the table is unchanged), "knn_sym"
 * (16-bit searches with k + 1 <= 16 and D <= 128: -1 by size (2^18 rows or more), 0 the list
 * tiles, s in 1..64 the symmetric threshold search with row bounds from every s-th row). */
int ncvis_tune(const char *key, double value);

/* ncvis::NCVis::fit_transform: X [N, D] and Y [N, d] are host, row-major float32. */
int ncvis_fit_transform(ncvis_handle *h, const float *X, size_t N, size_t D, float *Y);

/* Same with X and Y already in device memory (HBM). */
int ncvis_fit_transform_device(ncvis_handle *h, const float *dX, size_t N, size_t D, float *dY, void *stream);

/* Many independent data sets in one call, each fitted with this handle's parameters exactly as
 * ncvis_fit_transform_device fits it (the reference's fit_transform, src/ncvis.cpp:354-459, takes
 * one data set; this batched form has no counterpart there). dX[i] [N[i], D[i]] and dY[i] [N[i], d]
 * are device pointers. All data sets share each launch: a kNN search per data set, one symmetrize
 * over the concatenated tables of each neighbour count, then init and SGD launches that walk every
 * data set, so small data sets fill the GPU together. The work is captured once in a hipGraph and
 * replayed by later calls with the same pointers, sizes and parameters (option "graph" 0: queued
 * eagerly instead). n_streams is accepted for source compatibility and ignored. Runs on `stream`
 * (where the inputs were produced) and returns when every fit is done. An error names the first
 * failing data set; the handle's stage times are left unchanged. */
int ncvis_fit_transform_many(ncvis_handle *h, size_t count, const float *const *dX, const size_t *N, const size_t *D,
                             float *const *dY, int n_streams, void *stream);

/* Stage times of the last fit (milliseconds): [knn, symmetrize, init, optimize, total, knn_tiles,
 * knn_rerank, sgd_kernel] (sgd_kernel: the epoch kernel launches alone, summed; NaN for a fit
 * replayed from its captured graph, which records the stages but not every launch), the edge
 * count E, the final Q and the effective chain count. */
int ncvis_get_stats(const ncvis_handle *h, double *stage_ms8, int64_t *n_edges, float *final_Q, int64_t *chains);

/* SGD epoch mode of the last fit (no reference counterpart): 0 one exact chain with plain
 * stores (the reference at one thread), 1 exact chains with float atomics, 2 many chains with
 * device-scope atomics on Y, 3 many chains on per-XCD copies of Y merged after every sub-epoch;
 * launches: epoch kernel launches (epochs x sub-epochs), the count the sgd_kernel time of
 * ncvis_get_stats sums over. Either output may be NULL. */
int ncvis_get_sgd_mode(const ncvis_handle *h, int *mode, int *launches);

/* ---- stage entry points (device pointers) ---- */
const char *ncvis_get_last_error(void);

/* findKNN row format: d_inds [N, k] int32 / d_dists [N, k] float32, farthest first. */
int ncvis_knn_device(const float *dX, size_t N, size_t D, size_t k, int dist, int precision, int32_t *d_inds,
                     float *d_dists, void *stream);

/* Symmetrized CSR: d_row_ptr [N+1] int64, d_col / d_dist capacity 2*N*k. *n_edges receives E. */
int ncvis_symmetrize_device(const int32_t *d_inds, const float *d_dists, size_t N, size_t k, int64_t *d_row_ptr,
                            int32_t *d_col, float *d_dist, int64_t *n_edges, void *stream);

/* Random layout + n_init_epochs propagation epochs; n_threads = the reference's thread count
 * whose static schedule splits the random draws. The graph is validated first (row_ptr starts
 * at 0 and never decreases, columns lie in [0, N)); a bad graph returns NCVIS_EINVAL. */
int ncvis_init_embedding_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, float alpha,
                                int n_init_epochs, uint64_t seed, int64_t n_threads, float *dY, void *stream);

/* SGD epochs. Q is host in/out; q_trace (host, optional) receives Q after each epoch.
 * n_threads = the exact chain count (< 1: GPU-sized). n_noise entries must be below 2^32 - 1.
 * atomic_updates: 0 plain stores, 1 or -1 device atomics, 2 per-XCD copies of Y (d <= 4; the
 * copies are allocated here). 
 * The graph is validated first (row_ptr [N+1] from 0 to n_edges, never decreasing; columns in
 * [0, N)); a bad graph returns NCVIS_EINVAL. dY needs only float alignment. */
int ncvis_optimize_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t n_edges,
                          float *dY, float *Q, int n_epochs, const size_t *n_noise, float a, float b, float alpha,
                          float alpha_Q, uint64_t seed, int64_t n_threads, int atomic_updates, float *q_trace,
                          void *stream);

/* ---- kNN tile autotuning (no reference counterpart: the reference's HNSW has no tiles) ----
 * ncvis_autotune_knn times every admissible launch variant of the kNN candidate tiles for one
 * search shape (candidate split 1/2/4, the 16-bit LDS stage size, FP32 LDS DMA, and for 16-bit
 * searches of k + 1 <= 16 and D <= 128 the symmetric threshold search with row bounds from every
 * 4th, 8th or 16th row) on dX [N, D]
 * (device rows; NULL: uniform synthetic rows) and keeps the fastest in a process-wide table keyed
 * by device architecture, GEMM precision, padded width, list length and floor(log2 N). Every
 * variant gives the exact table on exactly scored data; more splits only add candidates to the
 * exact rerank. cache_path: NULL keeps the table in memory, "" writes it to
 * $NCVIS_AMD_TUNE_CACHE, any other string writes it to that JSON file. Outputs (may be NULL): the
 * chosen variant and its kNN time in ms. The table is read from $NCVIS_AMD_TUNE_CACHE at the first
 * search of a process; ncvis_load_tune_cache replaces it from a file (NULL or "": empties it).
 * Errors: ncvis_autotune_last_error. */
int ncvis_autotune_knn(const float *dX, size_t N, size_t D, size_t k, int dist, int precision, const char *cache_path,
                       int *split_out, int *stage_mult_out, int *dma_out, int *sym_out, double *ms_out);
int ncvis_load_tune_cache(const char *path);
int ncvis_tune_cache_size(void);

/* The kNN launch variant a search of this shape uses now (tuning table, knobs, size rules): the
 * candidate split, the 16-bit stage multiplier, FP32 LDS DMA and the symmetric search's
 * bounding-row stride (0: list tiles). No reference counterpart. */
int ncvis_knn_plan(size_t N, size_t D, size_t k, int dist, int precision, int *split_out, int *stage_mult_out,
                   int *dma_out, int *sym_out);
const char *ncvis_autotune_last_error(void);

/* ---- measurement (no reference counterpart) ----
 * The many-chain SGD epochs are bound by random 8-byte-row loads each followed by an update of
 * that row. ncvis_calibrate_rmw measures the device-atomic epoch's pattern: `threads` threads
 * (multiple of 256) each load `iters` pcg64-random rows of a table of `rows` 8-byte entries, once
 * read-only and once with a device-scope 64-bit compare-and-swap update; outputs rows per
 * second. ncvis_calibrate_rmw_copies measures the per-XCD-copy epoch's pattern: the table is
 * held once per XCD, each wave reads a random row of its XCD's copy, reads it again and stores
 * the update (workgroup scope, served by that XCD's L2). ncvis_calibrate_rmw_eighths measures the
 * epochs' pattern on permuted copies (option sgd_neg_slices; rows >= 8): each XCD's rows drawn
 * from its own eighth of the copy, each read once and stored. */
int ncvis_calibrate_rmw(size_t rows, int64_t threads, int iters, double *load_rows_per_s, double *cas_rows_per_s);
int ncvis_calibrate_rmw_copies(size_t rows, int64_t threads, int iters, double *rows_per_s);
int ncvis_calibrate_rmw_eighths(size_t rows, int64_t threads, int iters, double *rows_per_s);
const char *ncvis_calibrate_last_error(void);

/* ncvis_optimize_device without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results as ncvis_optimize_device;
 * its scratch (chain states, Q copies, the per-epoch Q buffer, the graph-check flag) lives in one
 * static device array per GPU (4.75 MiB, held until the process ends). Limits: at most 2^17
 * chains (n_threads < 1 picks the GPU-sized count, which stays within it by default) and at
 * most 2^16 epochs; larger requests return NCVIS_EINVAL before any GPU work. Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (other host threads, other streams)
 * race and must be serialised by the caller. */
int ncvis_optimize_device_noalloc(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col,
                                  int64_t n_edges, float *dY, float *Q, int n_epochs, const size_t *n_noise, float a,
                                  float b, float alpha, float alpha_Q, uint64_t seed, int64_t n_threads,
                                  int atomic_updates, float *q_trace, void *stream);

/* NCE loss over explicit negatives d_neg [E, n] int32. d_loss / d_gradQ: device double;
 * d_gradY: device float [N, d], accumulated into (may be NULL). The graph and the negatives are
 * validated first (indices in [0, N)); a bad index returns NCVIS_EINVAL. */
int ncvis_nce_loss_device(size_t N, size_t d, const int64_t *d_row_ptr, const int32_t *d_col, int64_t n_edges,
                          const int32_t *d_neg, int n, const float *dY, float Q, float a, float b, double *d_loss,
                          float *d_gradY, double *d_gradQ, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NCVIS_AMD_H */
