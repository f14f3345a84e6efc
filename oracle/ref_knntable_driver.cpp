// Harness around the reference's own src/knntable.cpp (compiled from /root/reference by
// oracle/Makefile into oracle/_ref/). TEST INFRASTRUCTURE ONLY: it lets the tests check the
// oracle's symmetrize restatement against the reference implementation itself.
// The kNN table is handed over row-major ([N, k] ids and distances, findKNN order); the
// symmetrized table is returned flattened as CSR in row order (= build_edges order,
// src/ncvis.cpp:155-173). Runs KNNTable::symmetrize (src/knntable.cpp:24-99) with
// n_threads OpenMP threads; n_threads = 1 gives the deterministic ascending-source order.
#include <omp.h>

#include "knntable.hpp"

extern "C" long ref_symmetrize(const long *inds, const float *dists, long N, long k, int n_threads, long *row_ptr,
                               long *col, float *dist_out) {
    omp_set_num_threads(n_threads);
    ncvis::KNNTable table(N, k);
    for (long i = 0; i < N; ++i) {
        for (long j = 0; j < k; ++j) {
            table.inds[i].push_back(inds[i * k + j]);
            table.dists[i].push_back(dists[i * k + j]);
        }
    }
    table.symmetrize();
    long out = 0;
    row_ptr[0] = 0;
    for (long i = 0; i < N; ++i) {
        for (size_t j = 0; j < table.inds[i].size(); ++j) {
            col[out] = table.inds[i][j];
            dist_out[out] = table.dists[i][j];
            ++out;
        }
        row_ptr[i + 1] = out;
    }
    return out;
}
