// ncvis::NCVis for MI355X: the reference's C++ class API (src/ncvis.hpp:26-93) with the fit path
// running on the GPU through the C ABI of libncvis_amd.so (include/ncvis_amd.h).
#ifndef NCVIS_AMD_NCVIS_HPP
#define NCVIS_AMD_NCVIS_HPP

#include <cstddef>
#include <string>
#include <utility>

struct ncvis_handle;

namespace ncvis {
typedef std::pair<size_t, size_t> Edge;  // src/ncvis.hpp:26
enum Distance {                          // src/ncvis.hpp:27-32 (same ordinals)
    squared_L2,
    inner_product,
    cosine_similarity,
    correlation
};

class NCVis {
   public:
    /*!
    @brief Same parameters and defaults as the reference constructor (src/ncvis.hpp:54).

    n_threads is the reference's OpenMP thread count. Each GPU SGD chain plays one such thread
    (its own pcg64(random_seed + id), a static chunk of the edges and a private Q copy). The
    chain count is GPU-sized by default; set_option("exact_threads", 1) runs exactly n_threads
    chains (n_threads = 1 then reproduces the reference at one thread). M and ef_construction
    are accepted and ignored: the neighbour search is exact. Throws std::runtime_error on
    invalid arguments or when no GPU is usable.
    */
    NCVis(size_t d = 2, size_t n_threads = 1, size_t n_neighbors = 30, size_t M = 16, size_t ef_construction = 200,
          size_t random_seed = 42, int n_epochs = 50, int n_init_epochs = 20, float a = 1., float b = 1.,
          float alpha = 1., float alpha_Q = 1., size_t *n_noise = nullptr,
          ncvis::Distance dist = ncvis::Distance::squared_L2);
    ~NCVis();
    NCVis(const NCVis &) = delete;
    NCVis &operator=(const NCVis &) = delete;

    /*!
    @brief Embeds X [N, D] (row-major float32, host memory) into Y [N, d] (src/ncvis.hpp:66).
    Throws std::runtime_error with the reference's messages for N == 0, D == 0 or Y == nullptr.
    */
    void fit_transform(const float *const X, size_t N, size_t D, float *Y);

    /*! @brief Library options (see ncvis_set_option): "knn_precision", "atomic_updates",
        "exact_threads", "chains". */
    void set_option(const std::string &key, double value);

    /*! @brief Stage times of the last fit in ms: knn, symmetrize, init, optimize, total, knn tiles,
        knn rerank, SGD epoch kernels; plus edge count, final Q and the chain count used. */
    void stats(double *stage_ms8, long long *n_edges, float *final_Q, long long *chains) const;

   private:
    ncvis_handle *h_;
};
}  // namespace ncvis

#endif
