// Float64 sampled Schur build + arrow factor (see fp64_sampled.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "fp64_dense.h"

namespace sdpx {

struct Fp64Leaf {
    int64_t h = 0, k = 0;
    const double *rinv = nullptr, *q = nullptr, *w = nullptr, *b = nullptr;  // device
    double *hbuf = nullptr;  // device k x k: leaf matrix, then its Cholesky factor
};

struct Fp64SampledTimes {
    double build_ms = 0, leaves_ms = 0, border_ms = 0;
};

// S (t x t, upper) = C - sum Y^T Y; border_factor = chol(-S) (lower).
void sampled_arrow_f64(const std::vector<Fp64Leaf> &leaves, int64_t t, const double *C, double *S,
                       double *border_factor, int *d_info, Fp64SampledTimes *times,
                       hipStream_t st, const PivotRule *rule = nullptr);

}  // namespace sdpx
