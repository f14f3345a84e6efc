// Calibration of the random read-modify-write rate that bounds the many-chain SGD epoch.
//
// Every SGD sample reads one random 8-byte row of Y (d = 2) and updates it with one device-scope
// 64-bit compare-and-swap (the far endpoint; the source row stays in registers). The roofline of
// that kernel is therefore the rate at which the GPU completes random load + CAS pairs on a table
// of Y's size, not HBM bandwidth: a CAS on a line shared by the XCDs' separate L2 caches is
// performed at the memory side. k_rmw_probe issues exactly that pattern (threads x iters random
// rows, pcg64-drawn, the same launch shape as the epoch kernel) and ncvis_calibrate_rmw reports
// rows per second for load-only and for load + CAS, so the measured epoch rate can be stated as a
// fraction of the pattern's own ceiling.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "common.h"
#include "kernels.h"
#include "ncvis_amd.h"

namespace ncvis_hip {

template <int CAS>
__global__ __launch_bounds__(256) void k_rmw_probe(unsigned long long *table, uint64_t rows, int iters,
                                                   float *sink) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    U128 g = pcg_seed(777 + (uint64_t)t);
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
        const uint64_t r = uniform_index(g, rows);
        unsigned long long *addr = table + r;
        unsigned long long expect = *addr;
        if (CAS) {
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

extern "C" {

const char *ncvis_calibrate_last_error(void) { return g_cal_error.c_str(); }

int ncvis_calibrate_rmw(size_t rows, int64_t threads, int iters, double *load_rows_per_s, double *cas_rows_per_s) {
    try {
        if (rows < 1 || threads < 256 || iters < 1) throw std::invalid_argument("[ncvis_amd] calibrate: bad sizes");
        hipStream_t st = nullptr;
        NCVIS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{st};
        DevBuf<unsigned long long> table(rows);
        DevBuf<float> sink(1);
        k_rmw_fill<<<cdiv((int64_t)rows, 256), 256, 0, st>>>(table.p, rows);
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
        double rate[2] = {0.0, 0.0};
        for (int cas = 0; cas < 2; ++cas) {
            for (int rep = 0; rep < 3; ++rep) {  // the first run warms up; the best of two is kept
                NCVIS_HIP_CHECK(hipEventRecord(e0, st));
                if (cas)
                    k_rmw_probe<1><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
                else
                    k_rmw_probe<0><<<grid, 256, 0, st>>>(table.p, rows, iters, sink.p);
                NCVIS_LAUNCH_CHECK();
                NCVIS_HIP_CHECK(hipEventRecord(e1, st));
                NCVIS_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                NCVIS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) rate[cas] = std::max(rate[cas], (double)(threads / 256 * 256) * iters / (ms * 1e-3));
            }
        }
        if (load_rows_per_s) *load_rows_per_s = rate[0];
        if (cas_rows_per_s) *cas_rows_per_s = rate[1];
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

}  // extern "C"
