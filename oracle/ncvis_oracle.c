/*
 * CPU oracle for the NCVis fit path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker (or as the timed CPU baseline). The product path
 * (ncvis_amd) never links, imports or calls it.
 *
 * This is a plain-C restatement of the reference algorithm, following:
 *   src/ncvis.cpp:32-70    preprocess                      -> orc_preprocess
 *   src/ncvis.cpp:128-153  findKNN row order semantics     -> orc_knn_bruteforce (exact search
 *                          stands in for hnswlib, whose source is absent)
 *   src/ncvis.cpp:72-153   buildKNN + findKNN with HNSW    -> orc_hnsw_knn (oracle/hnsw.c, the
 *                          published HNSW algorithm; the CPU baseline's kNN stage)
 *   src/knntable.cpp:24-99 KNNTable::symmetrize            -> orc_symmetrize
 *   src/ncvis.cpp:155-173  build_edges                     -> CSR produced by orc_symmetrize
 *   src/ncvis.cpp:175-182  d_sqr                           -> d_sqr
 *   src/ncvis.cpp:184-273  init_embedding                  -> orc_init_embedding
 *   src/ncvis.cpp:275-352  optimize                        -> orc_optimize / orc_sgd_sample
 *   src/ncvis.cpp:354-459  fit_transform orchestration     -> orc_fit_transform
 * plus the third-party arithmetic the reference calls:
 *   pcg64 (pcg-cpp setseq_xsl_rr_128_64, O'Neill 2014)    -> pcg64_seed / pcg64_next
 *   libstdc++ 11 uniform_int_distribution<size_t>         -> uniform_index (Lemire _S_nd,
 *                          bits/uniform_int_dist.h:240-330)
 *   libstdc++ 11 uniform_real_distribution<float>(0,1)    -> uniform_unit_float
 *                          (bits/random.tcc:3345-3380, bits/random.h:1864-1871)
 *   libstdc++ 11 std::sort (introsort)                    -> intro_sort_idx
 *                          (bits/stl_algo.h:1796-1960, bits/stl_heap.h:130-430)
 *
 * How each function is pinned (tests/test_oracle.py):
 *   pcg64_next, uniform_index, uniform_unit_float: golden vectors written by
 *     tests/golden/make_golden_rng.cpp (pcg-cpp header vendored in pyarrow + libstdc++).
 *   orc_symmetrize: the reference's own src/knntable.cpp compiled into oracle/_ref, on
 *     random and tie-heavy tables, and committed fixtures made from it.
 *   orc_preprocess: numpy float64 closed form.
 *   orc_knn_bruteforce: numpy float64 exact search (bit-exact on integer data).
 *   orc_init_embedding: scipy.sparse product for one propagation step, mean/std
 *     invariants, the uniform_unit_float golden stream and the libgomp static partition.
 *   orc_sgd_sample: float64 central finite differences of the NCE loss (orc_nce_loss),
 *     an independent softplus formulation.
 *   orc_optimize: closed-form decay of a single edge with n_noise = 0 (step schedule),
 *     and the sample body above. T > 1 (chain partition, per-thread pcg64(seed + t), the
 *     Q average of ncvis.cpp:342-349): with alpha = 0 the steps vanish and every chain is
 *     deterministic, so its Q trace is pinned against an independent restatement of the
 *     chunking, seeding and averaging in tests/test_oracle.py, against the mean of
 *     one-thread runs on the single chunks, and against orc_optimize_chains. The T > 1 run
 *     with alpha != 0 is Hogwild and nondeterministic in the reference itself.
 *   orc_optimize_chains: equal to orc_optimize at T = 1; at alpha = 0 equal to
 *     orc_optimize at T threads and to the restatement above.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ pcg64 */
#define PCG_MULT (((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL)
#define PCG_INC (((u128)6364136223846793005ULL << 64) | (u128)1442695040888963407ULL)

typedef struct {
    u128 state;
} pcg64_t;

/* engine(itype state): state_ = bump(state + increment()) */
static inline void pcg64_seed(pcg64_t *g, uint64_t seed) {
    g->state = ((u128)seed + PCG_INC) * PCG_MULT + PCG_INC;
}

/* output_previous = false for 128-bit state: advance, then XSL-RR of the new state */
static inline uint64_t pcg64_next(pcg64_t *g) {
    g->state = g->state * PCG_MULT + PCG_INC;
    u128 s = g->state;
    unsigned rot = (unsigned)(s >> 122);
    uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

/* uniform_int_distribution<size_t>(0, range-1) with a 64-bit engine: Lemire _S_nd */
static inline uint64_t uniform_index(pcg64_t *g, uint64_t range) {
    u128 prod = (u128)pcg64_next(g) * (u128)range;
    uint64_t low = (uint64_t)prod;
    if (low < range) {
        uint64_t thr = (0 - range) % range;
        while (low < thr) {
            prod = (u128)pcg64_next(g) * (u128)range;
            low = (uint64_t)prod;
        }
    }
    return (uint64_t)(prod >> 64);
}

/* uniform_real_distribution<float>(0,1): generate_canonical<float,24> with one 64-bit draw */
static inline float uniform_unit_float(pcg64_t *g) {
    float sum = (float)pcg64_next(g);
    float tmp = 18446744073709551616.0f; /* 2^64 */
    float r = sum / tmp;
    if (r >= 1.0f) r = nextafterf(1.0f, 0.0f);
    return r * (1.0f - 0.0f) + 0.0f;
}

void orc_pcg64_stream(uint64_t seed, uint64_t n, uint64_t *out) {
    pcg64_t g;
    pcg64_seed(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = pcg64_next(&g);
}

void orc_uniform_index_stream(uint64_t seed, uint64_t range, uint64_t n, uint64_t *out) {
    pcg64_t g;
    pcg64_seed(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = uniform_index(&g, range);
}

void orc_uniform_float_stream(uint64_t seed, uint64_t n, float *out) {
    pcg64_t g;
    pcg64_seed(&g, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = uniform_unit_float(&g);
}

/* libgomp static schedule for `#pragma omp for` without a chunk: thread t of T gets
 * [s0, e0) with q = n / T (+1 for the first n % T threads). */
void orc_static_chunk(int64_t n, int64_t T, int64_t t, int64_t *s0, int64_t *e0) {
    int64_t q = n / T, r = n % T;
    if (t < r) {
        r = 0;
        q++;
    }
    *s0 = q * t + r;
    *e0 = *s0 + q;
}

/* ------------------------------------------------------------- preprocess */
enum { DIST_L2 = 0, DIST_IP = 1, DIST_COSINE = 2, DIST_CORRELATION = 3 };

/* ncvis.cpp:32-70 */
static void preprocess_row(const float *x, int64_t D, int dist, float *out) {
    if (dist == DIST_CORRELATION) {
        float M = 0;
        for (int64_t i = 0; i < D; ++i) M += x[i];
        M /= D;
        for (int64_t i = 0; i < D; ++i) out[i] = x[i] - M;
    } else {
        for (int64_t i = 0; i < D; ++i) out[i] = x[i];
    }
    if (dist == DIST_CORRELATION || dist == DIST_COSINE) {
        float Nrm = 0;
        for (int64_t i = 0; i < D; ++i) Nrm += out[i] * out[i];
        if (Nrm != 0) {
            for (int64_t i = 0; i < D; ++i) out[i] /= sqrtf(Nrm);
        }
    }
}

void orc_preprocess(const float *X, int64_t N, int64_t D, int dist, float *out) {
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) preprocess_row(X + i * D, D, dist, out + i * D);
}

/* ------------------------------------------------------------------- kNN */
/* hnswlib distance semantics: L2Space = squared L2 (no sqrt); InnerProductSpace = 1 - <x,y>. */
static inline float pair_dist(const float *x, const float *y, int64_t D, int dist) {
    if (dist == DIST_L2) {
        float s = 0;
        for (int64_t i = 0; i < D; ++i) {
            float t = x[i] - y[i];
            s += t * t;
        }
        return s;
    }
    float s = 0;
    for (int64_t i = 0; i < D; ++i) s += x[i] * y[i];
    return 1.0f - s;
}

static inline int pair_less(float da, int64_t ia, float db, int64_t ib) {
    return da < db || (da == db && ia < ib);
}

/* Exact replacement of buildKNN + findKNN (ncvis.cpp:72-153): the k+1 smallest (dist, label)
 * pairs are found; the single smallest is dropped and the remaining k are stored farthest-first
 * (max-heap pop order of hnswlib's searchKnn result, ncvis.cpp:142-147). Xp is preprocessed. */
void orc_knn_bruteforce(const float *Xp, int64_t N, int64_t D, int64_t k, int dist, int n_threads,
                        int64_t *inds, float *dists) {
    int64_t kk = k + 1;
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        float *hd = (float *)malloc(sizeof(float) * kk);
        int64_t *hi = (int64_t *)malloc(sizeof(int64_t) * kk);
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < N; ++q) {
            /* sorted ascending list of the best kk */
            int64_t cnt = 0;
            const float *xq = Xp + q * D;
            for (int64_t c = 0; c < N; ++c) {
                float dv = pair_dist(xq, Xp + c * D, D, dist);
                if (cnt == kk && !pair_less(dv, c, hd[kk - 1], hi[kk - 1])) continue;
                int64_t p = (cnt < kk) ? cnt++ : kk - 1;
                while (p > 0 && pair_less(dv, c, hd[p - 1], hi[p - 1])) {
                    hd[p] = hd[p - 1];
                    hi[p] = hi[p - 1];
                    --p;
                }
                hd[p] = dv;
                hi[p] = c;
            }
            /* store the k largest, farthest first: positions kk-1 .. 1 */
            for (int64_t j = 0; j < k; ++j) {
                int64_t src = cnt - 1 - j;
                inds[q * k + j] = (src >= 1) ? hi[src] : -1;
                dists[q * k + j] = (src >= 1) ? hd[src] : INFINITY;
            }
        }
        free(hd);
        free(hi);
    }
}

/* -------------------------------------------- libstdc++ introsort over an index array */
/* comp(a, b) := key[a] < key[b] (knntable.cpp:82-83) */
#define KLESS(a, b) (key[(a)] < key[(b)])

static void push_heap_idx(int64_t *first, int64_t hole, int64_t top, int64_t value, const float *key) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && KLESS(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

static void adjust_heap_idx(int64_t *first, int64_t hole, int64_t len, int64_t value, const float *key) {
    const int64_t top = hole;
    int64_t second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (KLESS(first[second], first[second - 1])) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    push_heap_idx(first, hole, top, value, key);
}

static void make_heap_idx(int64_t *first, int64_t len, const float *key) {
    if (len < 2) return;
    int64_t parent = (len - 2) / 2;
    while (1) {
        adjust_heap_idx(first, parent, len, first[parent], key);
        if (parent == 0) return;
        parent--;
    }
}

static void heap_sort_idx(int64_t *first, int64_t len, const float *key) {
    /* __partial_sort(first, last, last): __heap_select (make_heap only) + __sort_heap */
    make_heap_idx(first, len, key);
    int64_t last = len;
    while (last > 1) {
        --last;
        int64_t value = first[last];
        first[last] = first[0];
        adjust_heap_idx(first, 0, last, value, key);
    }
}

static void move_median_to_first_idx(int64_t *a, int64_t r, int64_t x, int64_t y, int64_t z, const float *key) {
    int64_t pick;
    if (KLESS(a[x], a[y])) {
        if (KLESS(a[y], a[z])) pick = y;
        else if (KLESS(a[x], a[z])) pick = z;
        else pick = x;
    } else if (KLESS(a[x], a[z])) pick = x;
    else if (KLESS(a[y], a[z])) pick = z;
    else pick = y;
    int64_t t = a[r];
    a[r] = a[pick];
    a[pick] = t;
}

static int64_t unguarded_partition_idx(int64_t *a, int64_t first, int64_t last, int64_t pivot, const float *key) {
    while (1) {
        while (KLESS(a[first], a[pivot])) ++first;
        --last;
        while (KLESS(a[pivot], a[last])) --last;
        if (!(first < last)) return first;
        int64_t t = a[first];
        a[first] = a[last];
        a[last] = t;
        ++first;
    }
}

static void introsort_loop_idx(int64_t *a, int64_t first, int64_t last, int64_t depth, const float *key) {
    while (last - first > 16) {
        if (depth == 0) {
            heap_sort_idx(a + first, last - first, key);
            return;
        }
        --depth;
        int64_t mid = first + (last - first) / 2;
        move_median_to_first_idx(a, first, first + 1, mid, last - 1, key);
        int64_t cut = unguarded_partition_idx(a, first + 1, last, first, key);
        introsort_loop_idx(a, cut, last, depth, key);
        last = cut;
    }
}

static void unguarded_linear_insert_idx(int64_t *a, int64_t last, const float *key) {
    int64_t val = a[last];
    int64_t next = last - 1;
    while (key[val] < key[a[next]]) {
        a[last] = a[next];
        last = next;
        --next;
    }
    a[last] = val;
}

static void insertion_sort_idx(int64_t *a, int64_t first, int64_t last, const float *key) {
    if (first == last) return;
    for (int64_t i = first + 1; i != last; ++i) {
        if (KLESS(a[i], a[first])) {
            int64_t val = a[i];
            memmove(a + first + 1, a + first, sizeof(int64_t) * (size_t)(i - first));
            a[first] = val;
        } else {
            unguarded_linear_insert_idx(a, i, key);
        }
    }
}

static int64_t lg_floor(int64_t n) { return 63 - __builtin_clzll((unsigned long long)n); }

void orc_intro_sort_idx(int64_t *a, int64_t n, const float *key) {
    if (n == 0) return;
    introsort_loop_idx(a, 0, n, lg_floor(n) * 2, key);
    if (n > 16) {
        insertion_sort_idx(a, 0, 16, key);
        for (int64_t i = 16; i != n; ++i) unguarded_linear_insert_idx(a, i, key);
    } else {
        insertion_sort_idx(a, 0, n, key);
    }
}

/* ------------------------------------------------------------ symmetrize */
/* knntable.cpp:24-99 at one thread: incoming edges appended in ascending source order,
 * row := own k entries (findKNN order) ++ incoming, argsort by distance with std::sort,
 * then the in-place de-duplication whose comparison reads the row being overwritten
 * (knntable.cpp:85-93). Output is CSR (row i occupies [row_ptr[i], row_ptr[i+1])), which is
 * also the i-major edge order of build_edges (ncvis.cpp:155-173). Returns E. Output arrays
 * must hold 2*N*k entries. */
int64_t orc_symmetrize(const int64_t *inds, const float *dists, int64_t N, int64_t k, int64_t *row_ptr,
                       int64_t *col, float *dist_out) {
    int64_t *cnt = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    for (int64_t e = 0; e < N * k; ++e) cnt[inds[e] + 1]++;
    for (int64_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
    int64_t *fill = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    int64_t *in_src = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N * k + 1));
    float *in_dist = (float *)malloc(sizeof(float) * (size_t)(N * k + 1));
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j < k; ++j) {
            int64_t t = inds[i * k + j];
            int64_t p = cnt[t] + fill[t]++;
            in_src[p] = i;
            in_dist[p] = dists[i * k + j];
        }
    int64_t maxL = 0;
    for (int64_t i = 0; i < N; ++i) {
        int64_t L = k + (cnt[i + 1] - cnt[i]);
        if (L > maxL) maxL = L;
    }
    int64_t *r_ind = (int64_t *)malloc(sizeof(int64_t) * (size_t)(maxL + 1));
    float *r_dist = (float *)malloc(sizeof(float) * (size_t)(maxL + 1));
    int64_t *t_ind = (int64_t *)malloc(sizeof(int64_t) * (size_t)(maxL + 1));
    float *t_dist = (float *)malloc(sizeof(float) * (size_t)(maxL + 1));
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(maxL + 1));
    int64_t out = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        int64_t nin = cnt[i + 1] - cnt[i];
        int64_t L = k + nin;
        for (int64_t j = 0; j < k; ++j) {
            r_ind[j] = inds[i * k + j];
            r_dist[j] = dists[i * k + j];
        }
        for (int64_t j = 0; j < nin; ++j) {
            r_ind[k + j] = in_src[cnt[i] + j];
            r_dist[k + j] = in_dist[cnt[i] + j];
        }
        for (int64_t j = 0; j < L; ++j) idx[j] = j;
        memcpy(t_ind, r_ind, sizeof(int64_t) * (size_t)L);
        memcpy(t_dist, r_dist, sizeof(float) * (size_t)L);
        orc_intro_sort_idx(idx, L, r_dist);
        int64_t kk = 0;
        for (int64_t j = 0; j < L; ++j) {
            if (j == 0 || r_ind[idx[j]] != r_ind[idx[j - 1]]) {
                r_ind[kk] = t_ind[idx[j]];
                r_dist[kk] = t_dist[idx[j]];
                ++kk;
            }
        }
        for (int64_t j = 0; j < kk; ++j) {
            col[out + j] = r_ind[j];
            dist_out[out + j] = r_dist[j];
        }
        out += kk;
        row_ptr[i + 1] = out;
    }
    free(cnt);
    free(fill);
    free(in_src);
    free(in_dist);
    free(r_ind);
    free(r_dist);
    free(t_ind);
    free(t_dist);
    free(idx);
    return out;
}

/* ------------------------------------------------------------------- init */
/* ncvis.cpp:175-182 */
static inline float d_sqr(const float *x, const float *y, int64_t d) {
    float s = 0;
    for (int64_t i = 0; i < d; ++i) s += (x[i] - y[i]) * (x[i] - y[i]);
    return s;
}

/* ncvis.cpp:184-273, with the reference's OpenMP structure (n_threads threads). `src`/`dst`
 * is the flattened edge list of build_edges. */
void orc_init_embedding(int64_t N, int64_t d, int64_t E, const int64_t *src, const int64_t *dst, float alpha,
                        int n_init_epochs, uint64_t seed, int n_threads, float *Y) {
    float *Ys[2];
    Ys[0] = Y;
    Ys[1] = (float *)malloc(sizeof(float) * (size_t)(N * d));
    float *mean = (float *)malloc(sizeof(float) * (size_t)d);
    float *sigma = (float *)malloc(sizeof(float) * (size_t)d);
    omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        int id = omp_get_thread_num();
        pcg64_t g;
        pcg64_seed(&g, seed + (uint64_t)id);
#pragma omp for
        for (long long i = 0; i < N * d; ++i) Y[i] = uniform_unit_float(&g);
        for (int ep = 0; ep < n_init_epochs; ++ep) {
            float *Yo = Ys[ep % 2];
            float *Yn = Ys[(ep + 1) % 2];
#pragma omp for
            for (long long i = 0; i < N * d; ++i) Yn[i] = 0;
#pragma omp for
            for (long long e = 0; e < E; ++e) {
                int64_t a = src[e], b = dst[e];
                for (int64_t c = 0; c < d; ++c) Yn[a * d + c] += alpha * Yo[b * d + c];
            }
#pragma omp single
            for (int64_t c = 0; c < d; ++c) {
                mean[c] = 0;
                sigma[c] = 0;
            }
#pragma omp for
            for (long long i = 0; i < N; ++i)
                for (int64_t c = 0; c < d; ++c) {
#pragma omp atomic
                    mean[c] += Yn[i * d + c];
                }
#pragma omp single
            for (int64_t c = 0; c < d; ++c) mean[c] /= N;
#pragma omp for
            for (long long i = 0; i < N; ++i)
                for (int64_t c = 0; c < d; ++c) {
                    float delta2 = Yn[i * d + c] - mean[c];
                    delta2 *= delta2;
#pragma omp atomic
                    sigma[c] += delta2;
                }
#pragma omp single
            for (int64_t c = 0; c < d; ++c) {
                sigma[c] /= N;
                sigma[c] = sqrtf(sigma[c]);
            }
#pragma omp for
            for (long long i = 0; i < N; ++i)
                for (int64_t c = 0; c < d; ++c) Yn[i * d + c] = (Yn[i * d + c] - mean[c]) / sigma[c];
        }
        if (n_init_epochs % 2) {
#pragma omp for
            for (long long i = 0; i < N * d; ++i) Ys[0][i] = Ys[1][i];
        }
    }
    free(Ys[1]);
    free(mean);
    free(sigma);
}

/* ------------------------------------------------------------------ optimize */
typedef struct {
    float a, b, alpha_Q;
} sgd_params;

/* The body of ncvis.cpp:304-339 for one sample (id, other): returns the new Q_copy. */
static inline float sgd_sample(float *Y, int64_t d, int64_t id, int64_t other, int is_positive, uint64_t cur_noise,
                               float Q_copy, float step, sgd_params p) {
    float d2 = d_sqr(Y + id * d, Y + other * d, d);
    float Ph = 1 / (1 + p.a * powf(d2, p.b));
    float w = 1.;
    if (cur_noise != 0) {
        w = Ph / (cur_noise * expf(Q_copy));
        if (is_positive) {
            w = 1 / (1 + w);
        } else {
            w = -1 / (1 + 1 / w);
        }
        Q_copy -= w * p.alpha_Q;
        w = 2 * w * Ph * p.a * p.b * powf(d2, p.b - 1);
    }
    for (int64_t c = 0; c < d; ++c) {
        float dx = Y[other * d + c] - Y[id * d + c];
        dx = dx * w * step;
        if (dx > 4.) {
            dx = 4.;
        } else if (dx < -4.) {
            dx = -4.;
        }
        Y[id * d + c] += dx;
        Y[other * d + c] -= dx;
    }
    return Q_copy;
}

/* Exposed for the finite-difference check: one sample update, returns the new Q. */
float orc_sgd_sample(float *Y, int64_t d, int64_t id, int64_t other, int is_positive, uint64_t cur_noise, float Q,
                     float step, float a, float b, float alpha_Q) {
    sgd_params p = {a, b, alpha_Q};
    return sgd_sample(Y, d, id, other, is_positive, cur_noise, Q, step, p);
}

/* ncvis.cpp:275-352 with n_threads OpenMP threads (Hogwild for n_threads > 1). q_trace (may be
 * NULL) receives Q after each epoch. */
void orc_optimize(int64_t N, int64_t d, int64_t E, const int64_t *src, const int64_t *dst, float *Y, float *Q,
                  int n_epochs, const uint64_t *n_noise, float a, float b, float alpha, float alpha_Q, uint64_t seed,
                  int n_threads, float *q_trace) {
    float Q_cum = 0.;
    float Qs = *Q;
    sgd_params p = {a, b, alpha_Q};
    omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        int id_t = omp_get_thread_num();
        int nt = omp_get_num_threads();
        pcg64_t g;
        pcg64_seed(&g, seed + (uint64_t)id_t);
        for (int epoch = 0; epoch < n_epochs; ++epoch) {
            float step = alpha * (1 - (((float)epoch) / n_epochs) * (((float)epoch) / n_epochs));
            float Q_copy = Qs;
            uint64_t cur_noise = n_noise[epoch];
#pragma omp for nowait
            for (long long e = 0; e < E; ++e) {
                int64_t id = src[e];
                for (uint64_t j = 0; j < cur_noise + 1; ++j) {
                    int64_t other;
                    if (j == 0) {
                        other = dst[e];
                    } else {
                        do {
                            other = (int64_t)uniform_index(&g, (uint64_t)N);
                        } while (other == id);
                    }
                    Q_copy = sgd_sample(Y, d, id, other, j == 0, cur_noise, Q_copy, step, p);
                }
            }
#pragma omp atomic
            Q_cum += Q_copy;
#pragma omp barrier
#pragma omp single
            {
                Qs = Q_cum / nt;
                Q_cum = 0;
                if (q_trace) q_trace[epoch] = Qs;
            }
        }
    }
    *Q = Qs;
}

/* The same epochs with T chains run one after another inside each epoch (chain 0's whole static
 * chunk, then chain 1's, ...): chain t is OpenMP thread t of ncvis.cpp:279-349 -- its own
 * pcg64(seed + t) kept across epochs (:281-283), the libgomp static chunk of the edge loop
 * (:290-291), a private Q copy started from the epoch's Q (:288) -- and Q becomes the mean of the
 * T copies (:342-349), summed in chain order. This is one serialisation of the reference's T-thread
 * Hogwild run; with alpha = 0 every step is 0, Y never changes and each chain's Q copy no longer
 * depends on the others, so the T-thread run and this one agree up to the order of the float sum
 * of the T copies. q_trace (may be NULL) receives Q after each epoch. */
void orc_optimize_chains(int64_t N, int64_t d, int64_t E, const int64_t *src, const int64_t *dst, float *Y,
                         float *Q, int n_epochs, const uint64_t *n_noise, float a, float b, float alpha,
                         float alpha_Q, uint64_t seed, int64_t T, float *q_trace) {
    if (T < 1) T = 1;
    sgd_params p = {a, b, alpha_Q};
    pcg64_t *g = (pcg64_t *)malloc(sizeof(pcg64_t) * (size_t)T);
    for (int64_t t = 0; t < T; ++t) pcg64_seed(&g[t], seed + (uint64_t)t);
    float Qs = *Q;
    for (int epoch = 0; epoch < n_epochs; ++epoch) {
        float step = alpha * (1 - (((float)epoch) / n_epochs) * (((float)epoch) / n_epochs));
        uint64_t cur_noise = n_noise[epoch];
        float Q_cum = 0.;
        for (int64_t t = 0; t < T; ++t) {
            int64_t s0, e0;
            orc_static_chunk(E, T, t, &s0, &e0);
            float Q_copy = Qs;
            for (int64_t e = s0; e < e0; ++e) {
                int64_t id = src[e];
                for (uint64_t j = 0; j < cur_noise + 1; ++j) {
                    int64_t other;
                    if (j == 0) {
                        other = dst[e];
                    } else {
                        do {
                            other = (int64_t)uniform_index(&g[t], (uint64_t)N);
                        } while (other == id);
                    }
                    Q_copy = sgd_sample(Y, d, id, other, j == 0, cur_noise, Q_copy, step, p);
                }
            }
            Q_cum += Q_copy;
        }
        Qs = Q_cum / T;
        if (q_trace) q_trace[epoch] = Qs;
    }
    free(g);
    *Q = Qs;
}

/* ------------------------------------------------------------------ NCE loss */
/* The objective the update of ncvis.cpp:304-338 descends (derived; the reference has no loss
 * function): s = log Ph - Q - log n, Ph = 1/(1 + a d2^b);
 *   L = sum_e [ softplus(-s(i_e, j_e)) + sum_m softplus(s(i_e, neg_{e,m})) ].
 * Computed in float64. neg is [E, n] row-major. Gradients (if non-NULL) are accumulated
 * into dY [N, d] and *dQ. */
static inline double softplus(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

double orc_nce_loss(int64_t N, int64_t d, int64_t E, const int64_t *src, const int64_t *dst, const int64_t *neg,
                    int64_t n, const float *Y, double Q, double a, double b, double *dY, double *dQ) {
    (void)N;
    double L = 0, gQ = 0;
    double logn = log((double)n);
    for (int64_t e = 0; e < E; ++e) {
        int64_t i = src[e];
        for (int64_t m = 0; m <= n; ++m) {
            int64_t o = (m == 0) ? dst[e] : neg[e * n + (m - 1)];
            double d2 = 0;
            for (int64_t c = 0; c < d; ++c) {
                double t = (double)Y[i * d + c] - (double)Y[o * d + c];
                d2 += t * t;
            }
            double pw = pow(d2, b);
            double logPh = -log1p(a * pw);
            double s = logPh - Q - logn;
            /* dL/ds: positive -> -(1 - sigma(s)); negative -> sigma(s) */
            double sig = 1.0 / (1.0 + exp(-s));
            double g;
            if (m == 0) {
                L += softplus(-s);
                g = -(1.0 - sig);
            } else {
                L += softplus(s);
                g = sig;
            }
            gQ += -g;
            if (dY) {
                /* ds/dd2 = -a b d2^(b-1) / (1 + a d2^b) */
                double dsdd2 = (d2 > 0) ? -a * b * pw / d2 / (1.0 + a * pw) : 0.0;
                for (int64_t c = 0; c < d; ++c) {
                    double t = (double)Y[i * d + c] - (double)Y[o * d + c];
                    double gi = g * dsdd2 * 2.0 * t;
                    dY[i * d + c] += gi;
                    dY[o * d + c] -= gi;
                }
            }
        }
    }
    if (dQ) *dQ = gQ;
    return L;
}

/* ---------------------------------------------------------- fit_transform */
/* ncvis.cpp:354-459 with exact brute-force kNN in place of HNSW. Stage wall times (seconds)
 * are written to times[6] = {buildKNN (0: brute force has no build), findKNN, symmetrize,
 * build_edges, initialize, optimize}; the final Q to *Q_out. Returns 0 on success, -1 on bad
 * input. */
int orc_fit_transform(const float *X, int64_t N, int64_t D, int64_t d, int n_threads, int64_t n_neighbors,
                      uint64_t seed, int n_epochs, int n_init_epochs, float a, float b, float alpha, float alpha_Q,
                      const uint64_t *n_noise, int dist, float *Y, double *times, float *Q_out) {
    if (N == 0 || D == 0 || Y == NULL) return -1;
    double t0 = omp_get_wtime();
    float *Xp = (float *)malloc(sizeof(float) * (size_t)(N * D));
    omp_set_num_threads(n_threads);
    orc_preprocess(X, N, D, dist, Xp);
    int64_t k = (n_neighbors < N - 1) ? n_neighbors : (N - 1);
    k = (k > 0) ? k : 1;
    int64_t *inds = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N * k));
    float *dists = (float *)malloc(sizeof(float) * (size_t)(N * k));
    orc_knn_bruteforce(Xp, N, D, k, dist, n_threads, inds, dists);
    double t1 = omp_get_wtime();
    int64_t *row_ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * N * k));
    float *dout = (float *)malloc(sizeof(float) * (size_t)(2 * N * k));
    int64_t E = orc_symmetrize(inds, dists, N, k, row_ptr, col, dout);
    double t2 = omp_get_wtime();
    int64_t *src = (int64_t *)malloc(sizeof(int64_t) * (size_t)(E + 1));
    for (int64_t i = 0; i < N; ++i)
        for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) src[e] = i;
    double t3 = omp_get_wtime();
    float Q = 0.;
    float init_alpha = 1. / k;
    orc_init_embedding(N, d, E, src, col, init_alpha, n_init_epochs, seed, n_threads, Y);
    double t4 = omp_get_wtime();
    orc_optimize(N, d, E, src, col, Y, &Q, n_epochs, n_noise, a, b, alpha, alpha_Q, seed, n_threads, NULL);
    double t5 = omp_get_wtime();
    if (Q_out) *Q_out = Q;
    if (times) {
        times[0] = 0.0;
        times[1] = t1 - t0;
        times[2] = t2 - t1;
        times[3] = t3 - t2;
        times[4] = t4 - t3;
        times[5] = t5 - t4;
    }
    free(Xp);
    free(inds);
    free(dists);
    free(row_ptr);
    free(col);
    free(dout);
    free(src);
    return 0;
}
