// Command-line harness with the reference's arguments and synthetic data (src/main.cpp:9-44):
//   ncvis_cli [number of points] [number of neighbors] [number of threads] [maximum epochs]
//             [number of init epochs] [knn precision: fp16 | bf16 | fp32 (optional)]
//             [output file for Y as raw float32 (optional)]
// X is N x 784 float32 drawn as uniform_real<float>(0, 1) from pcg64(42), the embedding has
// d = 2, and the class is built as the reference does (M = 8, ef_construction = 200, seed 42,
// a = b = 1, default noise plan). Stage times of the fit are printed in the reference's
// "<stage>: <ms> ms" form (src/ncvis.cpp:370-439), then the whole call, then Y when N <= 5.
// The number of threads sets the SGD chain count only with NCVIS_EXACT_THREADS=1 in the
// environment (then 1 reproduces the reference at one thread); otherwise the GPU picks it.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <random>
#include <vector>

#include "ncvis.hpp"

namespace {

// pcg64 (PCG XSL-RR 128/64, O'Neill 2014, pcg-cpp's default multiplier, increment and seeding)
// as a UniformRandomBitGenerator, so the standard distribution draws what the reference draws.
struct Pcg64 {
    using result_type = uint64_t;
    using u128 = unsigned __int128;
    static constexpr u128 kMult = ((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL;
    static constexpr u128 kInc = ((u128)6364136223846793005ULL << 64) | (u128)1442695040888963407ULL;
    u128 state;
    explicit Pcg64(uint64_t seed) : state(((u128)seed + kInc) * kMult + kInc) {}
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~(result_type)0; }
    result_type operator()() {
        state = state * kMult + kInc;
        const unsigned rot = (unsigned)(state >> 122);
        const uint64_t x = (uint64_t)(state >> 64) ^ (uint64_t)state;
        return (x >> rot) | (x << ((64u - rot) & 63u));
    }
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        std::printf(
            "Usage: ncvis_cli [number of points] [number of neighbors] [number of threads] [maximum epochs] "
            "[number of init epochs] [knn precision: fp16|bf16|fp32] [output file]\n");
        return 1;
    }
    const long N = std::atol(argv[1]), D = 28 * 28, d = 2;
    const long n_threads = std::atol(argv[3]);
    if (N < 1) {
        std::fprintf(stderr, "number of points must be positive\n");
        return 1;
    }
    std::vector<float> X((size_t)N * D);
    Pcg64 pcg(42);
    std::uniform_real_distribution<float> gen_X(0, 1);
    for (long i = 0; i < N * D; ++i) X[(size_t)i] = gen_X(pcg);
    std::vector<float> Y((size_t)N * d);
    try {
        ncvis::NCVis vis(d, (size_t)n_threads, (size_t)std::atol(argv[2]), 8, 200, 42, std::atoi(argv[4]),
                         std::atoi(argv[5]));
        if (argc > 6) {
            const char *p = argv[6];
            const double code = !std::strcmp(p, "fp32") ? 0 : !std::strcmp(p, "bf16") ? 1 : !std::strcmp(p, "fp16") ? 2 : -1;
            if (code < 0) {
                std::fprintf(stderr, "knn precision must be fp16, bf16 or fp32\n");
                return 1;
            }
            vis.set_option("knn_precision", code);
        }
        const char *exact = std::getenv("NCVIS_EXACT_THREADS");
        if (exact && std::strcmp(exact, "0") != 0) vis.set_option("exact_threads", 1);
        const auto t0 = std::chrono::steady_clock::now();
        vis.fit_transform(X.data(), (size_t)N, (size_t)D, Y.data());
        const auto t1 = std::chrono::steady_clock::now();
        double ms[8];
        long long n_edges = 0, chains = 0;
        float Q = 0.f;
        vis.stats(ms, &n_edges, &Q, &chains);
        std::printf("knn: %.1f ms (candidate tiles %.1f ms, exact rerank %.1f ms)\n", ms[0], ms[5], ms[6]);
        std::printf("symmetrize: %.1f ms (with build_edges; E = %lld)\n", ms[1], n_edges);
        std::printf("initialize: %.1f ms\n", ms[2]);
        std::printf("optimize: %.1f ms (%lld chains, Q = %.4f)\n", ms[3], chains, Q);
        std::printf("fit_transform: %.1f ms (host buffers, including transfers)\n",
                    std::chrono::duration<double, std::milli>(t1 - t0).count());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    if (argc > 7) {
        FILE *f = std::fopen(argv[7], "wb");
        if (!f || std::fwrite(Y.data(), sizeof(float), Y.size(), f) != Y.size()) {
            std::fprintf(stderr, "cannot write %s\n", argv[7]);
            if (f) std::fclose(f);
            return 1;
        }
        std::fclose(f);
    }
    if (N <= 5) {
        std::printf("-----------------\n");
        for (long i = 0; i < N; ++i) {
            std::printf("[");
            for (long j = 0; j < d; ++j) std::printf("%6.2lf ", (double)Y[(size_t)(d * i + j)]);
            std::printf("]\n");
        }
        std::printf("-----------------\n");
    }
    return 0;
}
