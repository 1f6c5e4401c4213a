// Exact RNS product on device-resident values (see rns_gemm.hip).
#pragma once
#include <functional>
#include <vector>
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
    // fused subtraction: with sub set (an n x n upper product, matrix-core
    // CRT, 2, 4 or 8 limbs), entry (i, j) is not stored; instead
    // sub[i + j sub_ld] = RN(sub[i + j sub_ld] - RN(entry)) and *sub_done = 1.
    // Otherwise out is written as usual and *sub_done = 0.
    void *sub = nullptr;
    int64_t sub_ld = 0;
    int *sub_done = nullptr;
    // batched form (rns_gemm_batch): the scans already ran; their results
    // (int4 a row / column) and the 16 summary words (host-readable) are
    // read here instead of scanning and waiting
    const void *scan_a_in = nullptr, *scan_b_in = nullptr;
    const int64_t *summary = nullptr;
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
// Several products on the library stream with one host wait: every item's
// scans first, then one wait for all windows, then each item's encode,
// products and CRT (after(i) runs on the host once item i is queued, e.g.
// to queue work that consumes its output before the next item reuses the
// buffers).  Stops at the first item that does not return SDPX_GPU_OK and
// returns that code, with *failed its index.  Items must not set scratch
// or the batched-form fields.
int rns_gemm_batch(const std::vector<RnsGemmArgs> &items, const std::function<void(size_t)> &after,
                   size_t *failed);

}  // namespace sdpx
