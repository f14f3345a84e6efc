// Writes golden RNG vectors as JSON to stdout. Built and run by make_golden.py in this
// container only. It uses the pcg-cpp header that pyarrow vendors (namespace
// arrow_vendored::pcg_random) and libstdc++'s own distributions, i.e. exactly the pieces the
// reference combines in src/ncvis.cpp:195-196 and :281-283.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "arrow/vendored/pcg/pcg_random.hpp"

using pcg64 = arrow_vendored::pcg64;

static void raw(uint64_t seed, int n) {
    pcg64 g(seed);
    std::printf("\"pcg64_seed%llu\": [", (unsigned long long)seed);
    for (int i = 0; i < n; ++i) std::printf("%s\"%llu\"", i ? ", " : "", (unsigned long long)g());
    std::printf("],\n");
}

static void uint_draws(uint64_t seed, uint64_t N, int n, const char* tag) {
    pcg64 g(seed);
    std::uniform_int_distribution<size_t> gen(0, N - 1);
    std::printf("\"uniform_index_seed%llu_%s\": {\"range\": \"%llu\", \"draws\": [", (unsigned long long)seed, tag,
                (unsigned long long)N);
    for (int i = 0; i < n; ++i) std::printf("%s\"%llu\"", i ? ", " : "", (unsigned long long)gen(g));
    std::printf("]},\n");
}

static void float_draws(uint64_t seed, int n) {
    pcg64 g(seed);
    std::uniform_real_distribution<float> gen(0, 1);
    std::printf("\"uniform_float_bits_seed%llu\": [", (unsigned long long)seed);
    for (int i = 0; i < n; ++i) {
        float f = gen(g);
        uint32_t b;
        std::memcpy(&b, &f, 4);
        std::printf("%s%u", i ? ", " : "", b);
    }
    std::printf("],\n");
}

int main() {
    std::printf("{\n");
    raw(42, 16);
    raw(43, 16);
    raw(0, 16);
    uint_draws(42, 5620, 64, "n5620");
    uint_draws(42, 1000000, 64, "n1000000");
    uint_draws(7, 100, 64, "n100");
    // range 2^63 + 1 rejects about half of all draws: exercises the Lemire retry loop
    uint_draws(42, 9223372036854775809ULL, 64, "n2p63p1");
    float_draws(42, 64);
    float_draws(45, 64);
    std::printf("\"_generator\": \"tests/golden/make_golden_rng.cpp (pyarrow-vendored pcg-cpp + libstdc++ 11)\"\n}\n");
    return 0;
}
