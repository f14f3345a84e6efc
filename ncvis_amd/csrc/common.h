// Shared host/device helpers for the NCVis HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ncvis_hip {

// ---------------------------------------------------------------- errors
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define NCVIS_HIP_CHECK(expr)                                                                           \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            throw ::ncvis_hip::HipError(std::string("[ncvis_amd] HIP error '") + hipGetErrorString(_e) + \
                                        "' at " + __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr); \
    } while (0)

#define NCVIS_LAUNCH_CHECK() NCVIS_HIP_CHECK(hipGetLastError())

// Records e on st. On a stream under capture the record becomes an explicit event-record node of
// the graph being captured (after everything captured so far, before what follows), so every
// replay records e again and the stage times of a replayed fit are its own.
inline void record_event(hipEvent_t e, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    NCVIS_HIP_CHECK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusActive) {
        NCVIS_HIP_CHECK(hipEventRecord(e, st));
        return;
    }
    unsigned long long id = 0;
    hipGraph_t g = nullptr;
    const hipGraphNode_t *deps = nullptr;
    size_t ndeps = 0;
    NCVIS_HIP_CHECK(hipStreamGetCaptureInfo_v2(st, &cs, &id, &g, &deps, &ndeps));
    hipGraphNode_t node = nullptr;
    NCVIS_HIP_CHECK(hipGraphAddEventRecordNode(&node, g, deps, ndeps, e));
    NCVIS_HIP_CHECK(hipStreamUpdateCaptureDependencies(st, &node, 1, hipStreamSetCaptureDependencies));
}

// Stream for stream-ordered allocations made by this host thread (null: hipMalloc / hipFree).
// Batched fits run one data set per worker thread and stream; hipFree would wait for every
// stream of the device and serialise them.
inline hipStream_t &alloc_stream() {
    static thread_local hipStream_t s = nullptr;
    return s;
}

// Device memory owned by a host object (freed on scope exit).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t owner = nullptr;  // stream of a stream-ordered allocation
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), owner(o.owner) {
        o.p = nullptr;
        o.n = 0;
        o.owner = nullptr;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owner = o.owner;
            o.p = nullptr;
            o.n = 0;
            o.owner = nullptr;
        }
        return *this;
    }
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        owner = alloc_stream();
        if (owner)
            NCVIS_HIP_CHECK(hipMallocAsync((void **)&p, sizeof(T) * count, owner));
        else
            NCVIS_HIP_CHECK(hipMalloc((void **)&p, sizeof(T) * count));
    }
    void release() {
        if (p) {
            if (owner)
                (void)hipFreeAsync(p, owner);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        owner = nullptr;
    }
    ~DevBuf() { release(); }
};

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------ pcg64 on device
// pcg64 = pcg-cpp setseq_xsl_rr_128_64 with the default 128-bit multiplier and increment.
// Seeding follows the engine(itype) constructor: state = bump(seed + increment).
struct U128 {
    uint64_t lo, hi;
};

__host__ __device__ inline U128 mul128(U128 a, U128 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
#else
    uint64_t hi = (uint64_t)(((unsigned __int128)a.lo * b.lo) >> 64) + a.lo * b.hi + a.hi * b.lo;
#endif
    return {a.lo * b.lo, hi};
}

__host__ __device__ inline U128 add128(U128 a, U128 b) {
    uint64_t lo = a.lo + b.lo;
    return {lo, a.hi + b.hi + (lo < a.lo ? 1u : 0u)};
}

__host__ __device__ constexpr U128 pcg_mult() { return {4865540595714422341ULL, 2549297995355413924ULL}; }
__host__ __device__ constexpr U128 pcg_inc() { return {1442695040888963407ULL, 6364136223846793005ULL}; }

__host__ __device__ inline U128 pcg_seed(uint64_t seed) {
    return add128(mul128(add128(U128{seed, 0}, pcg_inc()), pcg_mult()), pcg_inc());
}

__host__ __device__ inline uint64_t pcg_next(U128 &s) {
    s = add128(mul128(s, pcg_mult()), pcg_inc());
    unsigned rot = (unsigned)(s.hi >> 58);
    uint64_t x = s.hi ^ s.lo;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// Jump the LCG ahead by `delta` steps (Brown, "Random number generation with arbitrary
// strides", 1994): O(log delta) 128-bit multiply-adds.
__host__ __device__ inline U128 pcg_advance(U128 s, uint64_t delta) {
    U128 acc_mult{1, 0}, acc_plus{0, 0};
    U128 cur_mult = pcg_mult(), cur_plus = pcg_inc();
    while (delta) {
        if (delta & 1) {
            acc_mult = mul128(acc_mult, cur_mult);
            acc_plus = add128(mul128(acc_plus, cur_mult), cur_plus);
        }
        cur_plus = mul128(add128(cur_mult, U128{1, 0}), cur_plus);
        cur_mult = mul128(cur_mult, cur_mult);
        delta >>= 1;
    }
    return add128(mul128(acc_mult, s), acc_plus);
}

// uniform_int_distribution<size_t>(0, range-1) on a 64-bit engine: Lemire's nearly
// divisionless method, the form libstdc++ 11 uses (_S_nd with a 128-bit product).
__device__ inline uint64_t uniform_index(U128 &s, uint64_t range) {
    uint64_t x = pcg_next(s);
    uint64_t low = x * range;
    uint64_t hi = __umul64hi(x, range);
    if (low < range) {
        uint64_t thr = (0 - range) % range;
        while (low < thr) {
            x = pcg_next(s);
            low = x * range;
            hi = __umul64hi(x, range);
        }
    }
    return hi;
}

// uniform_real_distribution<float>(0, 1): one 64-bit draw, float(u) / 2^64, clamped below 1.
__device__ inline float uniform_unit_float(U128 &s) {
    float f = (float)pcg_next(s);
    float r = f * 5.42101086242752217e-20f;  // 2^-64, exact power of two
    if (r >= 1.0f) r = __uint_as_float(0x3f7fffffu);
    return r;
}

// libgomp static schedule of `#pragma omp for` (no chunk): thread t of T owns [s0, e0).
__host__ __device__ inline void static_chunk(int64_t n, int64_t T, int64_t t, int64_t &s0, int64_t &e0) {
    int64_t q = n / T, r = n % T;
    if (t < r) {
        r = 0;
        q++;
    }
    s0 = q * t + r;
    e0 = s0 + q;
}

enum Distance : int { DIST_L2 = 0, DIST_IP = 1, DIST_COSINE = 2, DIST_CORRELATION = 3 };
enum Precision : int { PREC_FP32 = 0, PREC_BF16 = 1, PREC_FP16 = 2 };

}  // namespace ncvis_hip
