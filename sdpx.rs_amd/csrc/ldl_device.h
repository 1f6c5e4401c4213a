// Device-resident multiple-precision LDL^T and sampled arrow factor (see
// ldl_device.cpp).  All pointers are device pointers.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "ldl.h"
#include "sampled.h"

namespace sdpx {

struct SampledLeafDev {
    size_t h = 0, k = 0;
    const uint8_t *rinv = nullptr, *q = nullptr, *w = nullptr, *b = nullptr;
};

// 1 ok, 0 failed pivot, -1 not handled on the GPU (rerun on the host)
int ldl_factor_device(int limbs, uint8_t *d_l, size_t ld, size_t n, const int8_t *signs,
                      size_t nsigns, const void *reg_eps, const void *reg_delta, uint8_t *d_dinv,
                      LdlStats &stats, bool recursive);

// L D L^T x = b for nrhs columns of b (n x nrhs, ldb) with a factor from
// ldl_factor_device; 1 ok, -1 not handled on the GPU.
int ldl_solve_device(int limbs, const uint8_t *d_l, size_t ldl, size_t n, const uint8_t *d_dinv,
                     uint8_t *d_b, size_t ldb, size_t nrhs);

// leaf factors / D^-1 / Y land in contiguous arrays in leaf order
// (sum k^2, sum k, sum k*t values)
int sampled_arrow_device(int limbs, const std::vector<SampledLeafDev> &leaves, size_t t,
                         uint8_t *d_l_all, uint8_t *d_dinv_all, uint8_t *d_y_all, uint8_t *d_border,
                         const int8_t *border_signs, size_t nborder_signs, uint8_t *d_border_dinv,
                         SampledStats &st);

// Kernel-category times (ms) and event-pair counts of this thread's device
// pipelines while profiling is on: 0 ldl_dots panels, 1 triangular-solve
// bases, 2 forward rows, 3 exact products; reset clears them.
void panel_phase_times(double *ms, int64_t *launches, bool reset);

// ArrowLDLSolver::solve_many (arrow.rs:3793-3980) on the factors of
// sampled_arrow_device: x (n x nrhs, ldx, n = sum k + t; leaf rows in leaf
// order, then the border rows) holds b on entry, x on return.  1 ok, -1
// not handled on the GPU (a product declined, or an order above 8192).
int sampled_arrow_solve_device(int limbs, const std::vector<size_t> &ks, size_t t, const uint8_t *d_l_all,
                               const uint8_t *d_dinv_all, const uint8_t *d_y_all, const uint8_t *d_border,
                               const uint8_t *d_border_dinv, size_t nrhs, uint8_t *d_x, size_t ldx);

// The sampled Schur build alone (SampledSchurWorkspace::update_with_pool and
// entry, sampled/mod.rs:1836-2143, 1973-1988): leaf i's H (K x K, lower
// triangle, zeros above) at record sum_{j<i} K_j^2 of d_h_all, the same
// values sampled_arrow_device factors.  1 ok, -1 a product declined.
int sampled_schur_device(int limbs, const std::vector<SampledLeafDev> &leaves, uint8_t *d_h_all);

// Iterative-refinement settings (host records of the solve's precision):
// the reference's iterative_refinement_{reltol, abstol, stop_ratio, max_iter}
struct RefineSettings {
    const uint8_t *reltol = nullptr, *abstol = nullptr, *stop_ratio = nullptr;
    uint32_t max_iter = 0;
};

// DirectLDLKKTSolver::solve_many with stationary refinement (direct/
// solver.rs:470-520, refine 1046-1129) on the sampled arrow KKT
// K = [blockdiag(H_i) B; B^T C]: x = K^-1 b by sampled_arrow_solve_device,
// then per column the reference's loop, every residual RN(b - K x) exact and
// rounded once per row (ExactRows::residual, 1486-1670).  d_h_all as from
// sampled_schur_device, d_coupling[i] leaf i's B (K_i x t), d_c the border
// block C in its upper triangle (t x t), the factors as sampled_arrow_device
// left them.  b (n x nrhs, ldb) is read, x (ldx) written; ok[c] (host) the
// column's finiteness, norms (host, nrhs records) the infinity norm of the
// returned x's residual, refinements the correction solves summed over the
// columns.  1 ok, -1 a product declined (x unspecified).
int sampled_kkt_solve_device(int limbs, const std::vector<size_t> &ks, size_t t, const uint8_t *d_h_all,
                             const std::vector<const uint8_t *> &d_coupling, const uint8_t *d_c,
                             const uint8_t *d_l_all, const uint8_t *d_dinv_all, const uint8_t *d_y_all,
                             const uint8_t *d_border, const uint8_t *d_border_dinv, const RefineSettings &rs,
                             size_t nrhs, const uint8_t *d_b, size_t ldb, uint8_t *d_x, size_t ldx, int32_t *ok,
                             uint8_t *norms, int64_t *refinements);

}  // namespace sdpx
