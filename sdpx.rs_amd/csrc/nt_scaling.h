// Batched binary64 Nesterov-Todd scaling of PSD blocks (see nt_scaling.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sdpx {

// One block: order n, offset of its svec in s/z, of its n x n matrices in
// r/rinv/g, and of its n singular values.
struct NtBlock {
    int64_t n, svec_off, mat_off, vec_off;
};

constexpr int64_t NT_MAX_ORDER = 256;

// Global workspace for the blocks of order above 64 (three n x n matrices
// each; the smaller blocks work in LDS).
size_t nt_scaling_workspace_bytes(const std::vector<int64_t> &dims);

// Blocks of order <= 64 (d_small, device, nsmall of them, largest small_max)
// and the larger ones (d_large): status per block 0 ok, 1 S not positive
// definite, 2 Z not positive definite, 3 the Jacobi sweeps did not converge.
void nt_scaling_f64(const std::vector<int64_t> &dims, const NtBlock *d_small, int nsmall, int64_t small_max,
                    const NtBlock *d_large, int nlarge, int64_t large_max, const double *s_all, const double *z_all,
                    double *r_all, double *rinv_all, double *lam_all, double *g_all, double *work,
                    int *d_status_small, int *d_status_large, int max_sweeps, hipStream_t st);

}  // namespace sdpx
