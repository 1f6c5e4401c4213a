// Multiple-precision dense LDL^T (see ldl.cpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace sdpx {

struct LdlOptions {
    int threads = 1;
    bool gpu_products = true;
    double gpu_min_work = 32.0 * 32 * 32;  // m*n*k below this takes host dots
    bool reg = false;                      // regularization (eps, delta)
    const uint8_t *reg_eps = nullptr;
    const uint8_t *reg_delta = nullptr;
};

struct LdlStats {
    // declined: device panel paths that declined (the host panels ran);
    // retried: recursive factors that failed and were redone by exact dots
    int64_t gpu_products = 0, host_products = 0, declined = 0, regularized = 0, base_rows = 0;
    int64_t retried = 0;
    double gpu_mp_macs = 0;
};

std::vector<size_t> envelope(int limbs, const uint8_t *l, size_t ld, size_t n);

// Factor the lower triangle of l (n x n, ld) in place: unit lower L below the
// diagonal (ones on it), D^-1 into dinv.  recursive: ldl_rec (the reference's
// factor_rec, used from order 256) with the exact-dots retry of
// factor_loaded when it fails, else ldl_dots (factor_dots).  Returns false
// on a failed pivot (the reference's Err).
bool ldl_factor(int limbs, uint8_t *l, size_t ld, size_t n, const int8_t *signs, size_t nsigns,
                const LdlOptions &opt, uint8_t *dinv, LdlStats &stats, bool recursive);

}  // namespace sdpx
