// Sampled Schur build + arrow factor (see sampled.cpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "ldl.h"

namespace sdpx {

struct SampledLeafIn {
    size_t h = 0, k = 0;          // basis rows (block side), samples
    const uint8_t *rinv = nullptr;  // h x h, column-major
    const uint8_t *q = nullptr;     // h x k basis, column-major
    const uint8_t *w = nullptr;     // k weights
    const uint8_t *b = nullptr;     // k x t coupling to the border, column-major
};

struct SampledLeafOut {
    std::vector<uint8_t> l, dinv, y;  // factor (k x k), D^-1, Y = L^-1 B (k x t)
};

struct SampledStats {
    // declined: the device pipeline declined and the host panels ran;
    // retried: border factors redone by exact dots after a failed recursive one
    int64_t gpu_products = 0, host_products = 0, regularized = 0, failed_leaf = -1;
    int64_t declined = 0, retried = 0;
    double build_s = 0, leaves_s = 0, border_s = 0;
};

// border: t x t, column-major, upper triangle holds the border block C on
// entry; on success the lower triangle holds its unit LDL^T factor.
bool sampled_arrow_factor(int limbs, const std::vector<SampledLeafIn> &leaves, size_t t,
                          uint8_t *border, const int8_t *border_signs, size_t nborder_signs,
                          const LdlOptions &opt, std::vector<SampledLeafOut> &out,
                          uint8_t *border_dinv, SampledStats &st);

// The exact product used by the factor (GPU RNS above a size threshold).
void exact_product_public(int limbs, const LdlOptions &opt, char ta, char tb, size_t m, size_t n,
                          size_t k, const uint8_t *a, size_t lda, const uint8_t *b, size_t ldb,
                          uint8_t *c, bool packed, LdlStats &stats);

}  // namespace sdpx
