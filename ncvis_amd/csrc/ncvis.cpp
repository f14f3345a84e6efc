// ncvis::NCVis over the C ABI (include/ncvis_amd.h). Mirrors src/ncvis.cpp:12-30 (construction
// and destruction) and src/ncvis.cpp:354-459 (fit_transform); the work itself runs in
// libncvis_amd.so.
#include "ncvis.hpp"

#include <cstdint>
#include <stdexcept>

#include "ncvis_amd.h"

ncvis::NCVis::NCVis(size_t d, size_t n_threads, size_t n_neighbors, size_t M, size_t ef_construction,
                    size_t random_seed, int n_epochs, int n_init_epochs, float a, float b, float alpha, float alpha_Q,
                    size_t *n_noise, ncvis::Distance dist)
    : h_(nullptr) {
    int rc = ncvis_create(d, n_threads, n_neighbors, M, ef_construction, random_seed, n_epochs, n_init_epochs, a, b,
                          alpha, alpha_Q, n_noise, (int)dist, &h_);
    if (rc != 0) throw std::runtime_error(ncvis_get_last_error());
}

ncvis::NCVis::~NCVis() {
    ncvis_destroy(h_);
    h_ = nullptr;
}

void ncvis::NCVis::fit_transform(const float *const X, size_t N, size_t D, float *Y) {
    int rc = ncvis_fit_transform(h_, X, N, D, Y);
    if (rc != 0) throw std::runtime_error(ncvis_last_error(h_));
}

void ncvis::NCVis::set_option(const std::string &key, double value) {
    int rc = ncvis_set_option(h_, key.c_str(), value);
    if (rc != 0) throw std::runtime_error(ncvis_last_error(h_));
}

void ncvis::NCVis::stats(double *stage_ms8, long long *n_edges, float *final_Q, long long *chains) const {
    int64_t e = 0, c = 0;
    ncvis_get_stats(h_, stage_ms8, &e, final_Q, &c);
    if (n_edges) *n_edges = (long long)e;
    if (chains) *chains = (long long)c;
}
