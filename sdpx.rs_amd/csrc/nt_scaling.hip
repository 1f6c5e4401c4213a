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
// pairs of a step at once.  Blocks of order <= 64 run one workgroup each
// with their three working matrices in LDS; larger blocks take the steps
// below, spread over the chip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "common.h"
#include "fp64_dense.h"
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
                      double *rinv_all, double *lam_all, double *g_all, int *status, int max_sweeps) {
    extern __shared__ double lds[];
    __shared__ double red[4];
    __shared__ int perm[256];
    __shared__ double lam[256];
    const NtBlock blk = blocks[blockIdx.x];
    const int n = (int)blk.n;
    double *A = lds, *B = lds + n * n, *C = lds + 2 * n * n;  // L1; L2 then V; L2^T L1 then L1^-T V
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

// Blocks of order above 64: the same steps spread over the chip, each
// dense step on the FP64 MFMA routines (potrf_f64, gemm_f64,
// trsm_lower_t_f64) and each round-robin Jacobi step one launch of n/2
// workgroups (one column pair each, one row a thread).
__global__ void nt_zero_upper(double *A, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * n && e % n < e / n) A[e] = 0.0;
}

__global__ void nt_identity(double *V, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * n) V[e] = (e % n == e / n) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(NT_THREADS)
    nt_jacobi_step(double *C, double *V, int n, int st, int np, unsigned long long *off) {
    __shared__ double red[3][4];
    int p, q;
    rr_pair(st, blockIdx.x, np, p, q);
    if (q >= n) return;
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    double cp = 0.0, cq = 0.0;
    if (r < n) {
        cp = C[r + p * n];
        cq = C[r + q * n];
    }
    double al = cp * cp, be = cq * cq, ga = cp * cq;
    for (int o = 32; o > 0; o >>= 1) {
        al += __shfl_xor(al, o);
        be += __shfl_xor(be, o);
        ga += __shfl_xor(ga, o);
    }
    if (lane == 0) {
        red[0][wave] = al;
        red[1][wave] = be;
        red[2][wave] = ga;
    }
    __syncthreads();
    al = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    be = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    ga = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    if (!(al > 0.0 && be > 0.0)) return;
    const double nrm = sqrt(al) * sqrt(be), tol = DBL_EPSILON * n;
    if (!(fabs(ga) > tol * nrm)) return;
    if (r == 0) atomicMax(off, (unsigned long long)__double_as_longlong(fabs(ga) / nrm));
    const double zeta = (be - al) / (2.0 * ga);
    const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
    if (r < n) {
        C[r + p * n] = c * cp - sn * cq;
        C[r + q * n] = sn * cp + c * cq;
        const double vp = V[r + p * n], vq = V[r + q * n];
        V[r + p * n] = c * vp - sn * vq;
        V[r + q * n] = sn * vp + c * vq;
    }
}

// singular values (column norms of the rotated C), descending order
__global__ void __launch_bounds__(NT_THREADS) nt_values(const double *C, int n, double *lam, int *perm) {
    __shared__ double l[256];
    for (int j = threadIdx.x; j < n; j += NT_THREADS) {
        double acc = 0.0;
        for (int r = 0; r < n; r++) acc += C[r + j * n] * C[r + j * n];
        l[j] = sqrt(acc);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += NT_THREADS) {
        int rank = 0;
        for (int i = 0; i < n; i++) rank += (l[i] > l[j]) || (l[i] == l[j] && i < j);
        perm[rank] = j;
        lam[j] = l[j];
    }
}

// R[:, r] = (L1 V)[:, perm r] / sqrt(lambda), Rinv[r, :] = X[:, perm r]^T sqrt(lambda),
// lambda_out[r] = lambda[perm r]
__global__ void nt_finish(const double *LV, const double *X, const double *lam, const int *perm, int n, double *R,
                          double *Rinv, double *lam_out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int i = e % n, r = e / n;
    const int c = perm[r];
    R[i + r * n] = LV[i + c * n] * (1.0 / sqrt(lam[c]));
    const int rr = i, j = r;  // Rinv entry (rr, j)
    const int cr = perm[rr];
    Rinv[rr + j * n] = X[j + cr * n] * sqrt(lam[cr]);
    if (i == 0) lam_out[r] = lam[c];
}

}  // namespace

size_t nt_scaling_large_workspace_bytes(int64_t n) { return (size_t)(5 * n * n + n + 160) * sizeof(double); }

int nt_scaling_large_f64(int64_t n, const double *s, const double *z, double *R, double *Rinv, double *lam_out,
                         double *G, double *work, int max_sweeps, hipStream_t st) {
    const int nn = (int)n;
    double *S = work, *Z = S + n * n, *C = Z + n * n, *V = C + n * n, *X = V + n * n;
    double *lam = X + n * n;
    int *perm = (int *)(lam + n);
    unsigned long long *off = (unsigned long long *)(perm + 256);
    int *info = (int *)(off + 1);
    const unsigned grid = (unsigned)((n * n + 255) / 256);
    const int64_t sv = n * (n + 1) / 2;
    hipLaunchKernelGGL(svec_unpack, dim3((unsigned)((sv + 255) / 256)), dim3(256), 0, st, s, sv, n, (int64_t)1, S);
    hipLaunchKernelGGL(svec_unpack, dim3((unsigned)((sv + 255) / 256)), dim3(256), 0, st, z, sv, n, (int64_t)1, Z);
    HIP_CHECK(hipMemsetAsync(info, 0, 2 * sizeof(int), st));
    potrf_f64(S, n, n, info, nullptr, 0, st);
    potrf_f64(Z, n, n, info + 1, nullptr, 0, st);
    int hinfo[2];
    HIP_CHECK(hipMemcpyAsync(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (hinfo[0]) return 1;
    if (hinfo[1]) return 2;
    hipLaunchKernelGGL(nt_zero_upper, dim3(grid), dim3(256), 0, st, S, nn);
    hipLaunchKernelGGL(nt_zero_upper, dim3(grid), dim3(256), 0, st, Z, nn);
    {  // C = L2^T L1
        Fp64Gemm g;
        g.m = g.n = g.k = n;
        g.A = Z;
        g.lda = n;
        g.ta = true;
        g.B = S;
        g.ldb = n;
        g.C = C;
        g.ldc = n;
        gemm_f64(g, st);
    }
    hipLaunchKernelGGL(nt_identity, dim3(grid), dim3(256), 0, st, V, nn);
    const int npair = (nn + 1) / 2, np = 2 * npair;
    const double tol = DBL_EPSILON * n;
    bool converged = nn < 2;
    for (int sweep = 0; sweep < max_sweeps && !converged; sweep++) {
        HIP_CHECK(hipMemsetAsync(off, 0, sizeof *off, st));
        for (int step = 0; step < np - 1; step++)
            hipLaunchKernelGGL(nt_jacobi_step, dim3((unsigned)npair), dim3(NT_THREADS), 0, st, C, V, nn, step, np,
                               off);
        HIP_CHECK(hipGetLastError());
        unsigned long long h = 0;
        HIP_CHECK(hipMemcpyAsync(&h, off, sizeof h, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        double worst;
        std::memcpy(&worst, &h, sizeof worst);
        converged = worst <= tol;
    }
    if (!converged) return 3;
    hipLaunchKernelGGL(nt_values, dim3(1), dim3(NT_THREADS), 0, st, C, nn, lam, perm);
    {  // C = L1 V (the rotated C is no longer needed)
        Fp64Gemm g;
        g.m = g.n = g.k = n;
        g.A = S;
        g.lda = n;
        g.B = V;
        g.ldb = n;
        g.C = C;
        g.ldc = n;
        gemm_f64(g, st);
    }
    HIP_CHECK(hipMemcpyAsync(X, V, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    trsm_lower_t_f64(S, n, n, X, n, n, st);  // X = L1^-T V
    hipLaunchKernelGGL(nt_finish, dim3(grid), dim3(256), 0, st, C, X, lam, perm, nn, R, Rinv, lam_out);
    {  // G = R R^T
        Fp64Gemm g;
        g.m = g.n = g.k = n;
        g.A = R;
        g.lda = n;
        g.B = R;
        g.ldb = n;
        g.tb = true;
        g.C = G;
        g.ldc = n;
        gemm_f64(g, st);
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}

void nt_scaling_f64(const NtBlock *d_blocks, int nblocks, int64_t nmax, const double *s_all, const double *z_all,
                    double *r_all, double *rinv_all, double *lam_all, double *g_all, int *d_status, int max_sweeps,
                    hipStream_t st) {
    if (nblocks <= 0) return;
    if (nmax > NT_LDS_MAX) throw Error(SDPX_GPU_INVALID, "batched NT scaling takes orders up to 64");
    const size_t lds = 3 * (size_t)nmax * nmax * sizeof(double);
    HIP_CHECK(hipFuncSetAttribute((const void *)nt_scaling_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)std::max<size_t>(lds, 65536)));
    hipLaunchKernelGGL(nt_scaling_kernel, dim3(nblocks), dim3(NT_THREADS), lds, st, d_blocks, s_all, z_all, r_all,
                       rinv_all, lam_all, g_all, d_status, max_sweeps);
    HIP_CHECK(hipGetLastError());
}

}  // namespace sdpx
