# Declarations of the ncvis::NCVis C++ class (csrc/ncvis.hpp), as wrapper/cncvis.pxd of the
# reference declares its class; fit_transform is declared `except +` so C++ errors become
# Python RuntimeError.
from libcpp.string cimport string

cdef extern from "ncvis.hpp" namespace "ncvis":
    cdef enum Distance:
        squared_L2,
        inner_product,
        cosine_similarity,
        correlation

cdef extern from "ncvis.hpp" namespace "ncvis":
    cdef cppclass NCVis:
        NCVis(size_t d, size_t n_threads, size_t n_neighbors, size_t M, size_t ef_construction, size_t random_seed, int n_epochs, int n_init_epochs, float a, float b, float alpha, float alpha_Q, size_t* n_noise, Distance dist) except +
        void fit_transform(const float *const X, size_t N, size_t D, float* Y) except +
        void set_option(const string& key, double value) except +
        void stats(double* stage_ms8, long long* n_edges, float* final_Q, long long* chains)
