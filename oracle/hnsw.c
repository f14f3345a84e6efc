/*
 * CPU HNSW kNN for the oracle -- TEST INFRASTRUCTURE ONLY (the timed CPU baseline's kNN stage
 * and a recall reference for approximate search). The product path never links or calls it.
 *
 * The reference builds its kNN table with hnswlib (src/ncvis.cpp:72-153): an L2Space or
 * InnerProductSpace, HierarchicalNSW(space, N, M, ef_construction, random_seed), row 0 inserted
 * first, rows 1..N-1 inserted in parallel, then searchKnn(x, k+1) per row with hnswlib's
 * default ef (10, raised to k+1), the k farthest of the k+1 results kept farthest first.
 * hnswlib's source is not part of the reference tree; this is a restatement of the published
 * algorithm (Malkov & Yashunin, "Efficient and robust approximate nearest neighbor search using
 * Hierarchical Navigable Small World graphs", IEEE TPAMI 2018: Algorithms 1-4, with the
 * neighbour-selection heuristic of Algorithm 4 and hnswlib's conventions M_max = M on upper
 * layers, M_max0 = 2M on layer 0, level multiplier 1/ln(M)).
 *
 * Parity: unpinned against hnswlib itself (absent). Pinned by tests/test_oracle.py against the
 * exact search (recall on clustered data), the findKNN row-order invariants (farthest first,
 * smallest of the k+1 results dropped) and determinism at one thread.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float d;
    int32_t id;
} hn_pair;

/* ------------------------------------------------------------------ binary heaps */
/* max-heap (largest d on top) and min-heap over hn_pair; capacity grows as needed */
typedef struct {
    hn_pair *a;
    int n, cap;
} hn_heap;

static void hh_init(hn_heap *h, int cap) {
    h->a = (hn_pair *)malloc(sizeof(hn_pair) * (size_t)cap);
    h->n = 0;
    h->cap = cap;
}
static void hh_free(hn_heap *h) { free(h->a); }
static void hh_grow(hn_heap *h) {
    if (h->n < h->cap) return;
    h->cap *= 2;
    h->a = (hn_pair *)realloc(h->a, sizeof(hn_pair) * (size_t)h->cap);
}
/* sgn = +1: max-heap; sgn = -1: min-heap */
static void hh_push(hn_heap *h, float d, int32_t id, int sgn) {
    hh_grow(h);
    int i = h->n++;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (sgn * (h->a[p].d - d) >= 0.f) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].d = d;
    h->a[i].id = id;
}
static hn_pair hh_pop(hn_heap *h, int sgn) {
    hn_pair top = h->a[0];
    hn_pair last = h->a[--h->n];
    int i = 0;
    for (;;) {
        int c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && sgn * (h->a[c + 1].d - h->a[c].d) > 0.f) ++c;
        if (sgn * (h->a[c].d - last.d) <= 0.f) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n > 0) h->a[i] = last;
    return top;
}

/* ------------------------------------------------------------------ index */
typedef struct {
    const float *X;
    int64_t N;
    int D, is_l2, M, maxM0, efc;
    int32_t *links0;    /* [N, maxM0] */
    int32_t *cnt0;      /* [N] */
    int8_t *level;      /* [N] */
    int32_t **links_up; /* per node: [level, M] (levels 1..level) */
    int32_t **cnt_up;   /* per node: [level] */
    omp_lock_t *locks;  /* per node */
    omp_lock_t glock;
    int32_t enter;
    int maxlevel;
} hn_index;

static inline float hn_dist(const hn_index *ix, const float *a, int64_t j) {
    const float *b = ix->X + j * ix->D;
    float s = 0.f;
    if (ix->is_l2) {
        for (int t = 0; t < ix->D; ++t) {
            float u = a[t] - b[t];
            s += u * u;
        }
        return s;
    }
    for (int t = 0; t < ix->D; ++t) s += a[t] * b[t];
    return 1.0f - s;
}

static inline int32_t *hn_links(const hn_index *ix, int32_t id, int lev, int32_t **cnt) {
    if (lev == 0) {
        *cnt = ix->cnt0 + id;
        return ix->links0 + (int64_t)id * ix->maxM0;
    }
    *cnt = ix->cnt_up[id] + (lev - 1);
    return ix->links_up[id] + (int64_t)(lev - 1) * ix->M;
}

/* visited marks: one epoch-tagged array per thread */
typedef struct {
    uint32_t *mark;
    uint32_t tag;
} hn_visited;

static void hv_next(hn_visited *v, int64_t N) {
    if (++v->tag == 0) {
        memset(v->mark, 0, sizeof(uint32_t) * (size_t)N);
        v->tag = 1;
    }
}

/* Algorithm 2 (SEARCH-LAYER): returns the ef closest found in `top` (max-heap). `lock` guards
 * link reads while other threads insert. */
static void hn_search_layer(const hn_index *ix, const float *q, int32_t ep, int ef, int lev, hn_visited *vis,
                            hn_heap *top, hn_heap *cand, int lock) {
    top->n = 0;
    cand->n = 0;
    hv_next(vis, ix->N);
    float d0 = hn_dist(ix, q, ep);
    hh_push(top, d0, ep, 1);
    hh_push(cand, d0, ep, -1);
    vis->mark[ep] = vis->tag;
    float bound = d0;
    int32_t nb[512];
    while (cand->n > 0) {
        hn_pair c = hh_pop(cand, -1);
        if (c.d > bound && top->n >= ef) break;
        int32_t *cntp;
        if (lock) omp_set_lock(&((hn_index *)ix)->locks[c.id]);
        int32_t *l = hn_links(ix, c.id, lev, &cntp);
        int n = *cntp;
        memcpy(nb, l, sizeof(int32_t) * (size_t)n);
        if (lock) omp_unset_lock(&((hn_index *)ix)->locks[c.id]);
        for (int t = 0; t < n; ++t) {
            int32_t e = nb[t];
            if (vis->mark[e] == vis->tag) continue;
            vis->mark[e] = vis->tag;
            float d = hn_dist(ix, q, e);
            if (top->n < ef || d < bound) {
                hh_push(cand, d, e, -1);
                hh_push(top, d, e, 1);
                if (top->n > ef) hh_pop(top, 1);
                bound = top->a[0].d;
            }
        }
    }
}

/* Algorithm 4 (SELECT-NEIGHBORS-HEURISTIC, no extension, no pruned-connection refill): walk the
 * candidates nearest first and keep e unless a kept neighbour r is closer to e than the query is.
 * `w` holds the candidates (any order) and is overwritten by the kept ones, nearest first;
 * returns their count. */
static int hn_heuristic(const hn_index *ix, hn_pair *w, int n, int m) {
    /* ascending (d, id) insertion sort: candidate lists are short (<= ef_construction) */
    for (int i = 1; i < n; ++i) {
        hn_pair x = w[i];
        int j = i;
        while (j > 0 && (w[j - 1].d > x.d || (w[j - 1].d == x.d && w[j - 1].id > x.id))) {
            w[j] = w[j - 1];
            --j;
        }
        w[j] = x;
    }
    if (n <= m) return n;
    int kept = 0;
    for (int i = 0; i < n && kept < m; ++i) {
        const float *xe = ix->X + (int64_t)w[i].id * ix->D;
        int good = 1;
        for (int r = 0; r < kept; ++r) {
            if (hn_dist(ix, xe, w[r].id) < w[i].d) {
                good = 0;
                break;
            }
        }
        if (good) w[kept++] = w[i];
    }
    return kept;
}

/* connect `cur` to the heuristic's pick from `top` at layer `lev`, and back; returns the nearest
 * picked neighbour (the entry point for the next layer down) */
static int32_t hn_connect(hn_index *ix, int32_t cur, hn_heap *top, int lev, hn_pair *scratch) {
    const int mmax = lev == 0 ? ix->maxM0 : ix->M;
    int n = top->n;
    memcpy(scratch, top->a, sizeof(hn_pair) * (size_t)n);
    int m = hn_heuristic(ix, scratch, n, ix->M);
    int32_t *cntp;
    omp_set_lock(&ix->locks[cur]);
    int32_t *l = hn_links(ix, cur, lev, &cntp);
    for (int t = 0; t < m; ++t) l[t] = scratch[t].id;
    *cntp = m;
    omp_unset_lock(&ix->locks[cur]);
    hn_pair tmp[1025];
    for (int t = 0; t < m; ++t) {
        int32_t nbr = scratch[t].id;
        omp_set_lock(&ix->locks[nbr]);
        int32_t *nl = hn_links(ix, nbr, lev, &cntp);
        int c = *cntp;
        if (c < mmax) {
            nl[c] = cur;
            *cntp = c + 1;
        } else {
            const float *xn = ix->X + (int64_t)nbr * ix->D;
            for (int u = 0; u < c; ++u) {
                tmp[u].id = nl[u];
                tmp[u].d = hn_dist(ix, xn, nl[u]);
            }
            tmp[c].id = cur;
            tmp[c].d = hn_dist(ix, xn, cur);
            int kept = hn_heuristic(ix, tmp, c + 1, mmax);
            for (int u = 0; u < kept; ++u) nl[u] = tmp[u].id;
            *cntp = kept;
        }
        omp_unset_lock(&ix->locks[nbr]);
    }
    return m > 0 ? scratch[0].id : cur;
}

/* Algorithm 1 (INSERT) */
static void hn_insert(hn_index *ix, int32_t cur, hn_visited *vis, hn_heap *top, hn_heap *cand, hn_pair *scratch) {
    const float *q = ix->X + (int64_t)cur * ix->D;
    const int L = ix->level[cur];
    omp_set_lock(&ix->glock);
    int32_t ep = ix->enter;
    int maxl = ix->maxlevel;
    const int first = ep < 0;
    if (first) {
        ix->enter = cur;
        ix->maxlevel = L;
    }
    if (first || L <= maxl) omp_unset_lock(&ix->glock);
    if (first) return;
    /* greedy descent through the layers above L (ef = 1) */
    float dcur = hn_dist(ix, q, ep);
    for (int lev = maxl; lev > L; --lev) {
        int changed = 1;
        while (changed) {
            changed = 0;
            int32_t *cntp;
            omp_set_lock(&ix->locks[ep]);
            int32_t *l = hn_links(ix, ep, lev, &cntp);
            int n = *cntp;
            int32_t nb[64];
            memcpy(nb, l, sizeof(int32_t) * (size_t)n);
            omp_unset_lock(&ix->locks[ep]);
            for (int t = 0; t < n; ++t) {
                float d = hn_dist(ix, q, nb[t]);
                if (d < dcur) {
                    dcur = d;
                    ep = nb[t];
                    changed = 1;
                }
            }
        }
    }
    for (int lev = (L < maxl ? L : maxl); lev >= 0; --lev) {
        hn_search_layer(ix, q, ep, ix->efc, lev, vis, top, cand, 1);
        ep = hn_connect(ix, cur, top, lev, scratch);
    }
    if (L > maxl) {
        ix->enter = cur;
        ix->maxlevel = L;
        omp_unset_lock(&ix->glock);
    }
}

/* Algorithm 5 (K-NN-SEARCH) with ef = max(10, kq); results ascending (d, id) into out[kq] */
static int hn_search(const hn_index *ix, const float *q, int kq, hn_visited *vis, hn_heap *top, hn_heap *cand,
                     hn_pair *out) {
    int32_t ep = ix->enter;
    float dcur = hn_dist(ix, q, ep);
    for (int lev = ix->maxlevel; lev > 0; --lev) {
        int changed = 1;
        while (changed) {
            changed = 0;
            int32_t *cntp;
            int32_t *l = hn_links(ix, ep, lev, &cntp);
            int n = *cntp;
            for (int t = 0; t < n; ++t) {
                float d = hn_dist(ix, q, l[t]);
                if (d < dcur) {
                    dcur = d;
                    ep = l[t];
                    changed = 1;
                }
            }
        }
    }
    const int ef = kq > 10 ? kq : 10;
    hn_search_layer(ix, q, ep, ef, 0, vis, top, cand, 0);
    while (top->n > kq) hh_pop(top, 1);
    int n = top->n;
    for (int t = n - 1; t >= 0; --t) out[t] = hh_pop(top, 1);
    return n;
}

/* 64-bit splitmix draw for layer assignment (hnswlib draws -ln(U) * mult from its own engine) */
static inline uint64_t hn_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/*
 * buildKNN + findKNN (src/ncvis.cpp:72-153) on preprocessed rows Xp [N, D]: HNSW with M and
 * ef_construction, then per row the k+1 nearest found, smallest dropped, k written farthest
 * first. Rows with fewer than k+1 results stay -1 / inf (the reference leaves them empty).
 * times[0] = build seconds, times[1] = search seconds. Returns the number of short rows.
 */
long orc_hnsw_knn(const float *Xp, int64_t N, int D, int dist, int M, int efc, int k, uint64_t seed,
                  int n_threads, int64_t *inds, float *dists, double *times) {
    const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
    if (M < 2) M = 2;
    if (M > 32) M = 32;
    if (efc < M) efc = M;
    if (efc > 512) efc = 512;
    hn_index ix;
    memset(&ix, 0, sizeof(ix));
    ix.X = Xp;
    ix.N = N;
    ix.D = D;
    ix.is_l2 = (dist == 0);
    ix.M = M;
    ix.maxM0 = 2 * M;
    ix.efc = efc;
    ix.links0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * ix.maxM0);
    ix.cnt0 = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    ix.level = (int8_t *)malloc((size_t)N);
    ix.links_up = (int32_t **)calloc((size_t)N, sizeof(int32_t *));
    ix.cnt_up = (int32_t **)calloc((size_t)N, sizeof(int32_t *));
    ix.locks = (omp_lock_t *)malloc(sizeof(omp_lock_t) * (size_t)N);
    ix.enter = -1;
    ix.maxlevel = -1;
    omp_init_lock(&ix.glock);
    const double mult = 1.0 / log((double)M);
    for (int64_t i = 0; i < N; ++i) {
        double u = ((hn_mix(seed * 0x100000001B3ULL + (uint64_t)i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        int L = (int)(-log(u) * mult);
        if (L > 16) L = 16;
        ix.level[i] = (int8_t)L;
        if (L > 0) {
            ix.links_up[i] = (int32_t *)malloc(sizeof(int32_t) * (size_t)L * M);
            ix.cnt_up[i] = (int32_t *)calloc((size_t)L, sizeof(int32_t));
        }
        omp_init_lock(&ix.locks[i]);
    }
    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nt)
    {
        hn_visited vis;
        vis.mark = (uint32_t *)calloc((size_t)N, sizeof(uint32_t));
        vis.tag = 0;
        hn_heap top, cand;
        hh_init(&top, 2 * efc + 2);
        hh_init(&cand, 4 * efc + 16);
        hn_pair *scratch = (hn_pair *)malloc(sizeof(hn_pair) * (size_t)(efc + 2));
#pragma omp single
        hn_insert(&ix, 0, &vis, &top, &cand, scratch); /* row 0 first (ncvis.cpp:96-99) */
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 1; i < N; ++i) hn_insert(&ix, (int32_t)i, &vis, &top, &cand, scratch);
        free(scratch);
        hh_free(&top);
        hh_free(&cand);
        free(vis.mark);
    }
    double t1 = omp_get_wtime();
    long short_rows = 0;
#pragma omp parallel num_threads(nt) reduction(+ : short_rows)
    {
        hn_visited vis;
        vis.mark = (uint32_t *)calloc((size_t)N, sizeof(uint32_t));
        vis.tag = 0;
        hn_heap top, cand;
        hh_init(&top, 64);
        hh_init(&cand, 128);
        hn_pair *res = (hn_pair *)malloc(sizeof(hn_pair) * (size_t)(k + 2));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < N; ++i) {
            int n = hn_search(&ix, Xp + i * D, k + 1, &vis, &top, &cand, res);
            int64_t *oi = inds + i * k;
            float *od = dists + i * k;
            if (n < k + 1) {
                for (int j = 0; j < k; ++j) {
                    oi[j] = -1;
                    od[j] = INFINITY;
                }
                ++short_rows;
                continue;
            }
            /* res ascending: res[0] (the smallest, normally the row itself) is left out; the
             * row gets res[k], ..., res[1] (findKNN pops the max-heap k times) */
            for (int j = 0; j < k; ++j) {
                oi[j] = res[k - j].id;
                od[j] = res[k - j].d;
            }
        }
        free(res);
        hh_free(&top);
        hh_free(&cand);
        free(vis.mark);
    }
    double t2 = omp_get_wtime();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
    }
    for (int64_t i = 0; i < N; ++i) {
        omp_destroy_lock(&ix.locks[i]);
        free(ix.links_up[i]);
        free(ix.cnt_up[i]);
    }
    omp_destroy_lock(&ix.glock);
    free(ix.locks);
    free(ix.links_up);
    free(ix.cnt_up);
    free(ix.level);
    free(ix.cnt0);
    free(ix.links0);
    return short_rows;
}
