// Exact RNS product on device-resident values (see rns_gemm.hip).
#pragma once
#include <cstdint>

namespace sdpx {

struct RnsGemmArgs {
    int limbs;          // N: 64-bit limbs per mantissa
    char ta, tb;        // 'N' or 'T'
    int64_t m, n, k;
    const void *a;      // device values, column-major, leading dimension lda
    int64_t lda;
    const void *b;
    int64_t ldb;
    bool upper;         // only i <= j
    uint64_t out_len;   // packed upper when upper, m == n and out_len == n(n+1)/2
    void *out;          // device values, ldc = m unless packed
    // allocation-free form: every temporary in this device memory (null:
    // the library's workspace); more than scratch_bytes is a capacity error
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // expansion output (split_capacity > 0): instead of the rounded product,
    // chunk c (c < *split_used <= split_capacity) of every entry's exact value
    // at out + c m n records (full layout, ldc = m; upper entries only when
    // upper), each chunk P bits of the exact integer taken from the top, so
    // the chunks sum to the exact product; declined when more chunks would be
    // needed or the plan takes the VALU CRT
    int split_capacity = 0;
    int *split_used = nullptr;
};

// Size of the per-device array of the allocation-free exact product.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
// Looks up the current device's array address (once per device; called at
// device set-up), and returns it.
void rns_noalloc_setup();
void *rns_noalloc_scratch();

struct RnsStats {
    int moduli = 0, planes = 0;
    double bits = 0;
    int64_t width_a = 0, width_b = 0;
    // stage times (ms) when profiling is on: scan, encode, product, crt
    double stage_ms[4] = {0, 0, 0, 0};
    int64_t product_launches = 0;
};

// Record HIP events around the pipeline stages of later calls.
void rns_set_profiling(bool on);
bool rns_profiling();

// Totals over this thread's profiled calls since the last reset: stage ms
// (scan, encode, product, crt), algorithmic INT8 ops of the products, calls.
struct RnsTotals {
    double stage_ms[4] = {0, 0, 0, 0};
    double int8_ops = 0, mp_macs = 0;
    int64_t calls = 0;
};
RnsTotals &rns_totals();

// Runs on the library stream; the results are complete once the work queued
// so far on that stream is (the call waits only for its scan windows).
// Returns an SDPX_GPU_* code; throws sdpx::Error on device failures.
int rns_gemm_device(const RnsGemmArgs &args);
const RnsStats &rns_last_stats();

}  // namespace sdpx
