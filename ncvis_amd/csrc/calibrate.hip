// Calibration of the random read-modify-write rates that bound the many-chain SGD epochs.
//
// Every SGD sample reads one random 8-byte row of Y (d = 2) and updates it (the far endpoint; the
// source row stays in registers). With device atomics (mode 2) the update is a device-scope 64-bit
// compare-and-swap, performed at the memory side because the XCDs' L2 caches are not coherent with
// each other; on per-XCD copies of Y (mode 3) it is a second read and a plain store of the XCD's
// own copy, completed in that XCD's L2. The roofline of either kernel is the rate at which the GPU
// completes that pattern on a table of Y's size, not HBM bandwidth. k_rmw_probe issues exactly
// those patterns (threads x iters random rows, pcg64-drawn, the same launch shape as the epoch
// kernel); ncvis_calibrate_rmw reports rows per second for load-only and for load + CAS,
// ncvis_calibrate_rmw_copies for the per-XCD-copy update, so the measured epoch rate can be
// stated as a fraction of its pattern's own ceiling.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "common.h"
#include "kernels.h"
#include "ncvis_amd.h"

namespace ncvis_hip {

// MODE 0: load only; 1: load + device-scope 64-bit CAS on one table (the device-atomic epoch);
// 2: the per-XCD-copy epoch's update: the table holds kXcds copies, a wave works on its XCD's
// copy (hardware XCC id), reads the row, reads it again and stores the update, all at workgroup
// scope (served by the XCD's L2); 3: the epochs' pattern on permuted copies: the rows drawn from
// the eighth of the copy given by the XCD, each read once and stored (no second read).
template <int MODE>
__global__ __launch_bounds__(256) void k_rmw_probe(unsigned long long *table, uint64_t rows, int iters,
                                                   float *sink) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    U128 g = pcg_seed(777 + (uint64_t)t);
    float acc = 0.f;
    unsigned long long *tab = table;
    uint64_t lo = 0, range = rows;
    if (MODE >= 2) {
        const int x = __builtin_amdgcn_s_getreg((20 /* HW_REG_XCC_ID */) | (0 << 6) | (15 << 11)) & (kXcds - 1);
        tab = table + (uint64_t)x * rows;
        if (MODE == 3) {
            lo = rows * (uint64_t)x / kXcds;
            range = rows * (uint64_t)(x + 1) / kXcds - lo;
        }
    }
    for (int it = 0; it < iters; ++it) {
        const uint64_t r = lo + uniform_index(g, range);
        unsigned long long *addr = tab + r;
        if (MODE >= 2) {
            const unsigned long long first = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            acc += __builtin_bit_cast(float2, first).x;
            const unsigned long long cur =
                MODE == 3 ? first : __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            float2 v = __builtin_bit_cast(float2, cur);
            float2 nv = make_float2(__fsub_rn(v.x, 1e-7f), __fadd_rn(v.y, 1e-7f));
            __hip_atomic_store(addr, __builtin_bit_cast(unsigned long long, nv), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            continue;
        }
        unsigned long long expect = *addr;
        if (MODE == 1) {
            while (true) {
                float2 v = __builtin_bit_cast(float2, expect);
                float2 nv = make_float2(__fsub_rn(v.x, 1e-7f), __fadd_rn(v.y, 1e-7f));
                const unsigned long long prev = atomicCAS(addr, expect, __builtin_bit_cast(unsigned long long, nv));
                if (prev == expect) break;
                expect = prev;
            }
        } else {
            acc += __builtin_bit_cast(float2, expect).x;
        }
    }
    if (acc == 12345.f) *sink = acc;  // keeps the load-only reads alive
}

__global__ void k_rmw_fill(unsigned long long *table, uint64_t rows) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) table[i] = __builtin_bit_cast(unsigned long long, make_float2(1.f, -1.f));
}

}  // namespace ncvis_hip

using namespace ncvis_hip;

static thread_local std::string g_cal_error;

// Best of two timed runs (after one warm-up) of k_rmw_probe<MODE>: rows per second.
static double probe_rate(int mode, size_t rows, int64_t threads, int iters) {
    if (rows < 1 || threads < 256 || iters < 1) throw std::invalid_argument("[ncvis_amd] calibrate: bad sizes");
    hipStream_t st = nullptr;
    NCVIS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    if (mode == 3 && rows < (size_t)kXcds) throw std::invalid_argument("[ncvis_amd] calibrate: rows < 8");
    const size_t entries = rows * (mode >= 2 ? kXcds : 1);
    DevBuf<unsigned long long> table(entries);
    DevBuf<float> sink(1);
    k_rmw_fill<<<cdiv((int64_t)entries, 256), 256, 0, st>>>(table.p, entries);
    NCVIS_LAUNCH_CHECK();
    hipEvent_t e0, e1;
    NCVIS_HIP_CHECK(hipEventCreate(&e0));
    NCVIS_HIP_CHECK(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    const dim3 grid((unsigned)(threads / 256));
    double rate = 0.0;
    for (int rep = 0; rep < 3; ++rep) {  // the first run warms up; the best of two is kept
        NCVIS_HIP_CHECK(hipEventRecord(e0, st));
        if (mode == 0)
            k_rmw_probe<0><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
        else if (mode == 1)
            k_rmw_probe<1><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
        else if (mode == 2)
            k_rmw_probe<2><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
        else
            k_rmw_probe<3><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
        NCVIS_LAUNCH_CHECK();
        NCVIS_HIP_CHECK(hipEventRecord(e1, st));
        NCVIS_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        NCVIS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) rate = std::max(rate, (double)(threads / 256 * 256) * iters / (ms * 1e-3));
    }
    return rate;
}

template <typename F>
static int cal_guarded(F &&f) {
    try {
        f();
        g_cal_error.clear();
        return NCVIS_OK;
    } catch (const HipError &e) {
        g_cal_error = e.what();
        return NCVIS_EHIP;
    } catch (const std::exception &e) {
        g_cal_error = e.what();
        return NCVIS_EINVAL;
    }
}

extern "C" {

const char *ncvis_calibrate_last_error(void) { return g_cal_error.c_str(); }

int ncvis_calibrate_rmw(size_t rows, int64_t threads, int iters, double *load_rows_per_s, double *cas_rows_per_s) {
    return cal_guarded([&] {
        const double l = probe_rate(0, rows, threads, iters);
        const double c = probe_rate(1, rows, threads, iters);
        if (load_rows_per_s) *load_rows_per_s = l;
        if (cas_rows_per_s) *cas_rows_per_s = c;
    });
}

int ncvis_calibrate_rmw_copies(size_t rows, int64_t threads, int iters, double *rows_per_s) {
    return cal_guarded([&] {
        const double r = probe_rate(2, rows, threads, iters);
        if (rows_per_s) *rows_per_s = r;
    });
}

int ncvis_calibrate_rmw_eighths(size_t rows, int64_t threads, int iters, double *rows_per_s) {
    return cal_guarded([&] {
        const double r = probe_rate(3, rows, threads, iters);
        if (rows_per_s) *rows_per_s = r;
    });
}

}  // extern "C"
