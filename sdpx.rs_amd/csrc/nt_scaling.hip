// Nesterov-Todd scaling of PSD blocks in binary64, every block of a batch in
// one launch (one workgroup a block): PSDTriangleCone::update_scaling
// (crates/solver/src/solver/cones/psdtrianglecone.rs:311-430).
//
// Per block of order n with svec'd slack s and dual z:
//   S = smat(s), Z = smat(z)        (off-diagonals times 1/sqrt(2): svec_to_mat,
//                                     algebra/dense/matrix_math.rs:156-179)
//   L1 = chol(S), L2 = chol(Z)      (lower)
//   L2^T L1 = U diag(lambda) V^T    (singular values descending)
//   R    = L1 V diag(lambda)^-1/2
//   Rinv = diag(lambda)^1/2 (L1^-T V)^T, so Rinv R = I
//   G    = R R^T                     (the congruence X -> G X G of the cone)
// The singular value decomposition is one-sided Jacobi (Hestenes) on the
// columns of L2^T L1, which never forms (L2^T L1)^T (L2^T L1) and so, like
// the reference's direct SVD, does not square the condition number; the
// columns are rotated in parallel pairs by the round-robin ordering, all
// pairs of a step at once.  Blocks of order <= 64 keep their three working
// matrices in LDS, larger ones in a global workspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "common.h"
#include "nt_scaling.h"

namespace sdpx {
namespace {

constexpr int NT_THREADS = 256;
constexpr int NT_LDS_MAX = 64;  // largest order kept in LDS

__device__ inline void block_max(double &v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// In-place lower Cholesky of the n x n matrix A (ld n), strict upper set to
// zero; false when a pivot is not positive.
__device__ bool chol_lower(double *A, int n) {
    for (int k = 0; k < n; k++) {
        __syncthreads();
        const double d = A[k + k * n];
        if (!(d > 0.0)) return false;
        const double r = sqrt(d);
        for (int i = k + 1 + threadIdx.x; i < n; i += NT_THREADS) A[i + k * n] /= r;
        __syncthreads();
        if (threadIdx.x == 0) A[k + k * n] = r;
        const int m = n - k - 1;
        for (int e = threadIdx.x; e < m * m; e += NT_THREADS) {
            const int i = k + 1 + e % m, j = k + 1 + e / m;
            if (i >= j) A[i + j * n] -= A[i + k * n] * A[j + k * n];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) {
        const int i = e % n, j = e / n;
        if (i < j) A[e] = 0.0;
    }
    __syncthreads();
    return true;
}

// Round-robin pairing: step s of players 0 .. np-1 (np even), pair k.
__device__ inline void rr_pair(int s, int k, int np, int &p, int &q) {
    const int m = np - 1;
    if (k == 0) {
        p = m;
        q = s;
    } else {
        p = (s + k) % m;
        q = (s - k + m) % m;
    }
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
}

__global__ void __launch_bounds__(NT_THREADS)
    nt_scaling_kernel(const NtBlock *blocks, const double *s_all, const double *z_all, double *r_all,
                      double *rinv_all, double *lam_all, double *g_all, double *work, int64_t work_stride,
                      int *status, int max_sweeps, int use_lds) {
    extern __shared__ double lds[];
    __shared__ double red[4];
    __shared__ int perm[256];
    __shared__ double lam[256];
    const NtBlock blk = blocks[blockIdx.x];
    const int n = (int)blk.n;
    double *A, *B, *C;  // L1; L2 then V; L2^T L1 then L1^-T V
    if (use_lds) {
        A = lds;
        B = lds + n * n;
        C = lds + 2 * n * n;
    } else {
        A = work + (int64_t)blockIdx.x * work_stride;
        B = A + (int64_t)n * n;
        C = B + (int64_t)n * n;
    }
    const double *s = s_all + blk.svec_off, *z = z_all + blk.svec_off;
    const int sv = n * (n + 1) / 2;
    const double rs2 = M_SQRT1_2;
    for (int idx = threadIdx.x; idx < sv; idx += NT_THREADS) {
        int j = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > idx) j--;
        while ((j + 1) * (j + 2) / 2 <= idx) j++;
        const int i = idx - j * (j + 1) / 2;
        const double a = i == j ? s[idx] : s[idx] * rs2, b = i == j ? z[idx] : z[idx] * rs2;
        A[i + j * n] = a;
        A[j + i * n] = a;
        B[i + j * n] = b;
        B[j + i * n] = b;
    }
    if (!chol_lower(A, n)) {
        if (threadIdx.x == 0) status[blockIdx.x] = 1;
        return;
    }
    if (!chol_lower(B, n)) {
        if (threadIdx.x == 0) status[blockIdx.x] = 2;
        return;
    }
    // C = L2^T L1 (upper-times-lower: the sum runs over k >= max(i, j))
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) {
        const int i = e % n, j = e / n;
        double acc = 0.0;
        for (int k = i > j ? i : j; k < n; k++) acc += B[k + i * n] * A[k + j * n];
        C[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) B[e] = (e % n == e / n) ? 1.0 : 0.0;
    __syncthreads();
    // one-sided Jacobi: rotate column pairs of C (and V = B) until every
    // pair is orthogonal to |gamma| <= tol sqrt(alpha beta)
    const int npair = (n + 1) / 2, np = 2 * npair;
    int g = 1;
    while (g * 2 * npair <= NT_THREADS && g < 64) g *= 2;
    const int groups = NT_THREADS / g, lane = threadIdx.x % g, grp = threadIdx.x / g;
    const double tol = DBL_EPSILON * n;
    bool converged = n < 2;
    for (int sweep = 0; sweep < max_sweeps && !converged; sweep++) {
        double off = 0.0;
        for (int st = 0; st < np - 1; st++) {
            for (int k0 = 0; k0 < npair; k0 += groups) {
                const int k = k0 + grp;
                int p = 0, q = 0;
                const bool live = k < npair;
                if (live) rr_pair(st, k, np, p, q);
                const bool use = live && q < n;
                double al = 0.0, be = 0.0, ga = 0.0;
                if (use)
                    for (int r = lane; r < n; r += g) {
                        const double cp = C[r + p * n], cq = C[r + q * n];
                        al += cp * cp;
                        be += cq * cq;
                        ga += cp * cq;
                    }
                for (int o = g / 2; o > 0; o >>= 1) {  // groups are aligned lane ranges of one wave
                    al += __shfl_xor(al, o);
                    be += __shfl_xor(be, o);
                    ga += __shfl_xor(ga, o);
                }
                if (use && al > 0.0 && be > 0.0) {
                    const double nrm = sqrt(al) * sqrt(be);
                    if (fabs(ga) > tol * nrm) {
                        off = fmax(off, fabs(ga) / nrm);
                        const double zeta = (be - al) / (2.0 * ga);
                        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                        for (int r = lane; r < n; r += g) {
                            const double cp = C[r + p * n], cq = C[r + q * n];
                            C[r + p * n] = c * cp - sn * cq;
                            C[r + q * n] = sn * cp + c * cq;
                            const double vp = B[r + p * n], vq = B[r + q * n];
                            B[r + p * n] = c * vp - sn * vq;
                            B[r + q * n] = sn * vp + c * vq;
                        }
                    }
                }
            }
            __syncthreads();
        }
        block_max(off, red);
        converged = off <= tol;
    }
    if (!converged) {
        if (threadIdx.x == 0) status[blockIdx.x] = 3;
        return;
    }
    // singular values (column norms of C V-rotated), descending order
    for (int j = threadIdx.x; j < n; j += NT_THREADS) {
        double acc = 0.0;
        for (int r = 0; r < n; r++) acc += C[r + j * n] * C[r + j * n];
        lam[j] = sqrt(acc);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += NT_THREADS) {
        int rank = 0;
        for (int i = 0; i < n; i++) rank += (lam[i] > lam[j]) || (lam[i] == lam[j] && i < j);
        perm[rank] = j;
    }
    __syncthreads();
    double *R = r_all + blk.mat_off, *Rinv = rinv_all + blk.mat_off, *G = g_all + blk.mat_off;
    double *lo = lam_all + blk.vec_off;
    for (int r = threadIdx.x; r < n; r += NT_THREADS) lo[r] = lam[perm[r]];
    // R = L1 V diag(lambda)^-1/2 (the reference's R.rscale(1 / sqrt(lambda)))
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) {
        const int i = e % n, r = e / n, c = perm[r];
        double acc = 0.0;
        for (int k = 0; k <= i; k++) acc += A[i + k * n] * B[k + c * n];
        R[e] = acc * (1.0 / sqrt(lam[c]));
    }
    // X = L1^-T V, one column a thread (backward substitution), into C
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += NT_THREADS)
        for (int i = n - 1; i >= 0; i--) {
            double acc = B[i + c * n];
            for (int k = i + 1; k < n; k++) acc -= A[k + i * n] * C[k + c * n];
            C[i + c * n] = acc / A[i + i * n];
        }
    __syncthreads();
    // Rinv[r, j] = X[j, perm r] sqrt(lambda_perm r)
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) {
        const int r = e % n, j = e / n, c = perm[r];
        Rinv[e] = C[j + c * n] * sqrt(lam[c]);
    }
    __syncthreads();
    // G = R R^T, both triangles
    for (int e = threadIdx.x; e < n * n; e += NT_THREADS) {
        const int i = e % n, j = e / n;
        if (i > j) continue;
        double acc = 0.0;
        for (int r = 0; r < n; r++) acc += R[i + r * n] * R[j + r * n];
        G[i + j * n] = acc;
        G[j + i * n] = acc;
    }
    if (threadIdx.x == 0) status[blockIdx.x] = 0;
}

}  // namespace

size_t nt_scaling_workspace_bytes(const std::vector<int64_t> &dims) {
    int64_t big = 0, count = 0;
    for (int64_t n : dims)
        if (n > NT_LDS_MAX) {
            big = std::max(big, n);
            count++;
        }
    return (size_t)(3 * big * big * count) * sizeof(double);
}

void nt_scaling_f64(const std::vector<int64_t> &dims, const NtBlock *d_small, int nsmall, int64_t small_max,
                    const NtBlock *d_large, int nlarge, int64_t large_max, const double *s_all, const double *z_all,
                    double *r_all, double *rinv_all, double *lam_all, double *g_all, double *work,
                    int *d_status_small, int *d_status_large, int max_sweeps, hipStream_t st) {
    (void)dims;
    if (nsmall > 0) {
        const size_t lds = 3 * (size_t)small_max * small_max * sizeof(double);
        HIP_CHECK(hipFuncSetAttribute((const void *)nt_scaling_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)std::max<size_t>(lds, 65536)));
        hipLaunchKernelGGL(nt_scaling_kernel, dim3(nsmall), dim3(NT_THREADS), lds, st, d_small, s_all, z_all, r_all,
                           rinv_all, lam_all, g_all, nullptr, (int64_t)0, d_status_small, max_sweeps, 1);
        HIP_CHECK(hipGetLastError());
    }
    if (nlarge > 0) {
        hipLaunchKernelGGL(nt_scaling_kernel, dim3(nlarge), dim3(NT_THREADS), 0, st, d_large, s_all, z_all, r_all,
                           rinv_all, lam_all, g_all, work, (int64_t)(3 * large_max * large_max), d_status_large,
                           max_sweeps, 0);
        HIP_CHECK(hipGetLastError());
    }
}

}  // namespace sdpx
