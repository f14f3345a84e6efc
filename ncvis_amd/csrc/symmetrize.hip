// Bit-exact symmetrization of the kNN table on gfx950 (src/knntable.cpp:24-99 at one thread),
// flattened to CSR in row order, which is also build_edges' i-major edge order
// (src/ncvis.cpp:155-173).
//
//   k_count_in     incoming-edge counts (atomics)                      [HBM: N*k*4 read]
//   scan           in_ptr = exclusive prefix of counts (hipcub)
//   k_scatter_in   place flat edge ids e = i*k + j into target segments [N*k*(4+8) + scatter]
//   k_merge_rows   one thread per row: incoming ids sorted ascending (= the reference's
//                  one-thread append order), row = own k entries ++ incoming, argsort by distance
//                  with libstdc++'s introsort steps, then the in-place de-duplication whose
//                  comparison reads entries it has already overwritten (knntable.cpp:85-93)
//   scan           row_ptr = exclusive prefix of the kept lengths
//   k_compact      copy kept entries into the CSR arrays
//
// The argsort must reproduce std::sort's permutation (it is unstable for rows longer than 16
// and the de-duplication depends on it), so the introsort below follows libstdc++ 11
// (bits/stl_algo.h:1796-1960, bits/stl_heap.h:130-430) step for step. Disjoint sub-ranges are
// independent, so they are processed from an explicit stack (larger part deferred) instead of
// by recursion; that changes the visiting order only, not the result.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "common.h"
#include "kernels.h"

namespace ncvis_hip {

// ------------------------------------------------------------------ libstdc++ introsort
#define KLESS(x, y) (key[(x)] < key[(y)])

__device__ inline void dev_push_heap(int *f, int hole, int top, int value, const float *key) {
    int parent = (hole - 1) / 2;
    while (hole > top && KLESS(f[parent], value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

__device__ inline void dev_adjust_heap(int *f, int hole, int len, int value, const float *key) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (KLESS(f[second], f[second - 1])) second--;
        f[hole] = f[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        f[hole] = f[second - 1];
        hole = second - 1;
    }
    dev_push_heap(f, hole, top, value, key);
}

// __partial_sort(first, last, last) = __heap_select (a make_heap when middle == last) + __sort_heap
__device__ inline void dev_heap_sort(int *f, int len, const float *key) {
    if (len >= 2) {
        int parent = (len - 2) / 2;
        while (true) {
            dev_adjust_heap(f, parent, len, f[parent], key);
            if (parent == 0) break;
            parent--;
        }
    }
    int last = len;
    while (last > 1) {
        --last;
        int value = f[last];
        f[last] = f[0];
        dev_adjust_heap(f, 0, last, value, key);
    }
}

__device__ inline void dev_median_to_first(int *a, int r, int x, int y, int z, const float *key) {
    int pick;
    if (KLESS(a[x], a[y])) {
        if (KLESS(a[y], a[z]))
            pick = y;
        else if (KLESS(a[x], a[z]))
            pick = z;
        else
            pick = x;
    } else if (KLESS(a[x], a[z])) {
        pick = x;
    } else if (KLESS(a[y], a[z])) {
        pick = z;
    } else {
        pick = y;
    }
    int t = a[r];
    a[r] = a[pick];
    a[pick] = t;
}

__device__ inline int dev_unguarded_partition(int *a, int first, int last, int pivot, const float *key) {
    while (true) {
        while (KLESS(a[first], a[pivot])) ++first;
        --last;
        while (KLESS(a[pivot], a[last])) --last;
        if (!(first < last)) return first;
        int t = a[first];
        a[first] = a[last];
        a[last] = t;
        ++first;
    }
}

__device__ inline void dev_insertion_sort(int *a, int first, int last, const float *key) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        int val = a[i];
        if (key[val] < key[a[first]]) {
            for (int j = i; j > first; --j) a[j] = a[j - 1];
            a[first] = val;
        } else {
            int p = i, nx = i - 1;
            while (key[val] < key[a[nx]]) {
                a[p] = a[nx];
                p = nx;
                --nx;
            }
            a[p] = val;
        }
    }
}

__device__ inline int dev_lg(int n) { return 31 - __clz(n); }

// std::sort(a, a + n, [key](i, j) { return key[i] < key[j]; })
__device__ void dev_intro_sort(int *a, int n, const float *key) {
    if (n == 0) return;
    struct Range {
        int first, last, depth;
    };
    Range stack[40];
    int sp = 0;
    stack[sp++] = {0, n, 2 * dev_lg(n)};
    while (sp > 0) {
        Range r = stack[--sp];
        int first = r.first, last = r.last, depth = r.depth;
        while (last - first > 16) {
            if (depth == 0) {
                dev_heap_sort(a + first, last - first, key);
                break;
            }
            --depth;
            int mid = first + (last - first) / 2;
            dev_median_to_first(a, first, first + 1, mid, last - 1, key);
            int cut = dev_unguarded_partition(a, first + 1, last, first, key);
            // defer the larger part so the stack stays below log2(n) entries
            if (cut - first > last - cut) {
                stack[sp++] = {first, cut, depth};
                first = cut;
            } else {
                stack[sp++] = {cut, last, depth};
                last = cut;
            }
        }
    }
    if (n > 16) {
        dev_insertion_sort(a, 0, 16, key);
        for (int i = 16; i < n; ++i) {  // __unguarded_linear_insert (guarded for memory safety)
            int val = a[i];
            int p = i, nx = i - 1;
            while (nx >= 0 && key[val] < key[a[nx]]) {
                a[p] = a[nx];
                p = nx;
                --nx;
            }
            a[p] = val;
        }
    } else {
        dev_insertion_sort(a, 0, n, key);
    }
}
#undef KLESS

// ascending sort of distinct int64 values (incoming edge ids): insertion / Shell sort
__device__ void dev_sort_ids(int64_t *v, int64_t n) {
    if (n <= 32) {
        for (int64_t i = 1; i < n; ++i) {
            int64_t x = v[i];
            int64_t j = i;
            while (j > 0 && v[j - 1] > x) {
                v[j] = v[j - 1];
                --j;
            }
            v[j] = x;
        }
        return;
    }
    int64_t gap = 1;
    while (gap < n / 3) gap = 3 * gap + 1;
    for (; gap >= 1; gap /= 3) {
        for (int64_t i = gap; i < n; ++i) {
            int64_t x = v[i];
            int64_t j = i;
            while (j >= gap && v[j - gap] > x) {
                v[j] = v[j - gap];
                j -= gap;
            }
            v[j] = x;
        }
    }
}

// ------------------------------------------------------------------ kernels
template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t n, T v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_count_in(const int32_t *__restrict__ inds, int64_t Nk, int64_t N, unsigned long long *__restrict__ cnt,
                           int *__restrict__ err) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Nk) return;
    int32_t t = inds[e];
    if (t < 0 || t >= N) {
        atomicOr(err, kErrSymRange);
        return;
    }
    atomicAdd(&cnt[t], 1ull);
}

__global__ void k_scatter_in(const int32_t *__restrict__ inds, int64_t Nk, int64_t N, const int64_t *__restrict__ in_ptr,
                             unsigned long long *__restrict__ fill, int64_t *__restrict__ in_src) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Nk) return;
    int32_t t = inds[e];
    if (t < 0 || t >= N) return;  // flagged by k_count_in
    int64_t pos = in_ptr[t] + (int64_t)atomicAdd(&fill[t], 1ull);
    in_src[pos] = e;
}

__global__ void k_merge_rows(const int32_t *__restrict__ inds, const float *__restrict__ dists, int64_t N, int k,
                             const int64_t *__restrict__ in_ptr, int64_t *__restrict__ in_src,
                             int32_t *__restrict__ tind, float *__restrict__ tdist, int32_t *__restrict__ idx,
                             int32_t *__restrict__ alias, float *__restrict__ outd, int64_t *__restrict__ rowlen) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t ib = in_ptr[i];
    const int64_t nin = in_ptr[i + 1] - ib;
    const int L = (int)(k + nin);
    const int64_t base = i * k + ib;
    dev_sort_ids(in_src + ib, nin);
    int32_t *ti = tind + base;
    float *td = tdist + base;
    int32_t *ix = idx + base;
    int32_t *al = alias + base;
    float *od = outd + base;
    for (int j = 0; j < k; ++j) {
        ti[j] = inds[i * k + j];
        td[j] = dists[i * k + j];
    }
    for (int64_t j = 0; j < nin; ++j) {
        int64_t e = in_src[ib + j];
        ti[k + j] = (int32_t)(e / k);
        td[k + j] = dists[e];
    }
    for (int j = 0; j < L; ++j) {
        ix[j] = j;
        al[j] = ti[j];
    }
    dev_intro_sort(ix, L, td);
    int kk = 0;
    for (int j = 0; j < L; ++j) {
        if (j == 0 || al[ix[j]] != al[ix[j - 1]]) {
            al[kk] = ti[ix[j]];
            od[kk] = td[ix[j]];
            ++kk;
        }
    }
    rowlen[i] = kk;
}

__global__ void k_compact(int64_t N, int k, const int64_t *__restrict__ in_ptr, const int64_t *__restrict__ row_ptr,
                          const int32_t *__restrict__ alias, const float *__restrict__ outd,
                          int32_t *__restrict__ col, float *__restrict__ dist) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t base = i * k + in_ptr[i];
    const int64_t o = row_ptr[i];
    const int64_t n = row_ptr[i + 1] - o;
    for (int64_t j = 0; j < n; ++j) {
        col[o + j] = alias[base + j];
        dist[o + j] = outd[base + j];
    }
}

// exclusive prefix sum of `n` values into out[0..n], out[n] = total (no host synchronisation)
template <typename In>
static void exclusive_scan(const In *d_in, int64_t *d_out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t st) {
    k_fill<int64_t><<<1, 1, 0, st>>>(d_out, 1, 0);
    NCVIS_LAUNCH_CHECK();
    if (n == 0) return;
    size_t bytes = tmp_bytes;
    NCVIS_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, bytes, d_in, d_out + 1, (int)n, st));
}

// Workspace of symmetrize_enqueue, carved at 256-byte boundaries.
struct SymLayout {
    size_t cnt, in_ptr, in_src, tind, idx, alias, tdist, outd, rowlen, tmp, tmp_bytes, total;
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static SymLayout sym_layout(int64_t N, int k) {
    const int64_t Nk = N * (int64_t)k;
    const size_t cap = (size_t)(2 * Nk);
    size_t a = 0, b = 0;
    NCVIS_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, a, (const unsigned long long *)nullptr,
                                                     (int64_t *)nullptr, (int)N));
    NCVIS_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, b, (const int64_t *)nullptr, (int64_t *)nullptr,
                                                     (int)N));
    SymLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.cnt = take(sizeof(unsigned long long) * N);
    L.in_ptr = take(sizeof(int64_t) * (N + 1));
    L.in_src = take(sizeof(int64_t) * Nk);
    L.tind = take(sizeof(int32_t) * cap);
    L.idx = take(sizeof(int32_t) * cap);
    L.alias = take(sizeof(int32_t) * cap);
    L.tdist = take(sizeof(float) * cap);
    L.outd = take(sizeof(float) * cap);
    L.rowlen = take(sizeof(int64_t) * N);
    L.tmp_bytes = std::max<size_t>(std::max(a, b), 1);
    L.tmp = take(L.tmp_bytes);
    L.total = o;
    return L;
}

size_t symmetrize_workspace_bytes(int64_t N, int k) { return sym_layout(N, k).total; }

void symmetrize_enqueue(const int32_t *d_inds, const float *d_dists, int64_t N, int k, int64_t *d_row_ptr,
                        int32_t *d_col, float *d_dist, void *ws, int *d_err, hipStream_t st) {
    if (N < 1 || k < 1) throw HipError("[ncvis_amd] symmetrize: invalid sizes");
    if (N > INT32_MAX) throw HipError("[ncvis_amd] symmetrize: more than 2^31 points");
    const SymLayout L = sym_layout(N, k);
    unsigned char *w = (unsigned char *)ws;
    auto *cnt = (unsigned long long *)(w + L.cnt);
    auto *in_ptr = (int64_t *)(w + L.in_ptr);
    auto *in_src = (int64_t *)(w + L.in_src);
    auto *tind = (int32_t *)(w + L.tind);
    auto *idx = (int32_t *)(w + L.idx);
    auto *alias = (int32_t *)(w + L.alias);
    auto *tdist = (float *)(w + L.tdist);
    auto *outd = (float *)(w + L.outd);
    auto *rowlen = (int64_t *)(w + L.rowlen);
    void *tmp = w + L.tmp;
    const int64_t Nk = N * (int64_t)k;
    k_fill<unsigned long long><<<cdiv(N, 256), 256, 0, st>>>(cnt, N, 0ull);
    k_count_in<<<cdiv(Nk, 256), 256, 0, st>>>(d_inds, Nk, N, cnt, d_err);
    NCVIS_LAUNCH_CHECK();
    exclusive_scan(cnt, in_ptr, N, tmp, L.tmp_bytes, st);
    k_fill<unsigned long long><<<cdiv(N, 256), 256, 0, st>>>(cnt, N, 0ull);  // reused as fill counters
    k_scatter_in<<<cdiv(Nk, 256), 256, 0, st>>>(d_inds, Nk, N, in_ptr, cnt, in_src);
    NCVIS_LAUNCH_CHECK();
    k_merge_rows<<<cdiv(N, 128), 128, 0, st>>>(d_inds, d_dists, N, k, in_ptr, in_src, tind, tdist, idx, alias, outd,
                                                rowlen);
    NCVIS_LAUNCH_CHECK();
    exclusive_scan(rowlen, d_row_ptr, N, tmp, L.tmp_bytes, st);
    k_compact<<<cdiv(N, 256), 256, 0, st>>>(N, k, in_ptr, d_row_ptr, alias, outd, d_col, d_dist);
    NCVIS_LAUNCH_CHECK();
}

int64_t symmetrize_run(const int32_t *d_inds, const float *d_dists, int64_t N, int k, int64_t *d_row_ptr,
                       int32_t *d_col, float *d_dist, hipStream_t st) {
    DevBuf<unsigned char> ws(symmetrize_workspace_bytes(N, k));
    DevBuf<int> err(1);
    k_fill<int><<<1, 1, 0, st>>>(err.p, 1, 0);
    symmetrize_enqueue(d_inds, d_dists, N, k, d_row_ptr, d_col, d_dist, ws.p, err.p, st);
    int herr = 0;
    int64_t E = 0;
    NCVIS_HIP_CHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipMemcpyAsync(&E, d_row_ptr + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    NCVIS_HIP_CHECK(hipStreamSynchronize(st));
    if (herr & kErrSymRange) throw HipError("[ncvis_amd] symmetrize: neighbour index out of range");
    return E;
}

}  // namespace ncvis_hip
