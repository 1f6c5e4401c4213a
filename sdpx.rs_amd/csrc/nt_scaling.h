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

// Blocks of order <= 64 (d_blocks, device), one workgroup each: status per
// block 0 ok, 1 S not positive definite, 2 Z not positive definite, 3 the
// Jacobi sweeps did not converge.
void nt_scaling_f64(const NtBlock *d_blocks, int nblocks, int64_t nmax, const double *s_all, const double *z_all,
                    double *r_all, double *rinv_all, double *lam_all, double *g_all, int *d_status, int max_sweeps,
                    hipStream_t st);

// One block of order 65..256 spread over the chip (see nt_scaling.hip): the
// status code as above; work holds nt_scaling_large_workspace_bytes(n).
size_t nt_scaling_large_workspace_bytes(int64_t n);
int nt_scaling_large_f64(int64_t n, const double *s, const double *z, double *R, double *Rinv, double *lam_out,
                         double *G, double *work, int max_sweeps, hipStream_t st);

}  // namespace sdpx
