// Float64 form of the sampled Schur build + arrow factor, device-resident:
//   V = Rinv Q, Gram = V^T V, H_pq = w_p w_q Gram_pq^2      (FP64 MFMA GEMMs)
//   H = L L^T per leaf, Y = L^-1 B                            (blocked Cholesky, TRSM)
//   S = C - sum_leaves Y^T Y  (one FP64 MFMA GEMM over the stacked Y)
//   -S = L_b L_b^T                                            (border Cholesky)
// The same problem as the multiple-precision path (sampled/mod.rs:1836-2143;
// the binary64 dense quasi-definite elimination H = L L^T, Y = L^-1 B,
// S = C + Y^T Y of kkt/ldl/dense_block.rs:95-229), for the same-GPU
// comparison of the two precisions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "fp64_dense.h"
#include "fp64_sampled.h"

namespace sdpx {

// H (K x K, lower incl. diagonal) = w_p w_q G_pq^2 from the Gram's upper triangle.
__global__ void sampled_entries_f64(const double *G, int64_t K, const double *w, double *H) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K * K) return;
    int64_t p = idx % K, q = idx / K;
    if (p < q) return;
    double g = G[q + p * K];  // (q, p), q <= p, upper
    H[p + q * K] = w[p] * w[q] * 0.5 * (g * g + g * g);
}

__global__ void negate_upper_to_full(const double *S, int64_t t, double *out) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= t * t) return;
    int64_t i = idx % t, j = idx / t;
    out[i + j * t] = i <= j ? -S[i + j * t] : -S[j + i * t];
}

// Y = L^-1 B for the lower factor L (K x K) and B (K x t): forward half of
// the blocked solve.
void trsm_lower_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs, int64_t ldb,
                    hipStream_t s);

void sampled_arrow_f64(const std::vector<Fp64Leaf> &leaves, int64_t t, const double *C,
                       double *S, double *border_factor, int *d_info, Fp64SampledTimes *times,
                       hipStream_t st, const PivotRule *rule) {
    Workspace &ws = device().ws;
    int64_t ktot = 0, kmax = 0, hmax = 0;
    for (auto &l : leaves) {
        ktot += l.k;
        kmax = std::max(kmax, l.k);
        hmax = std::max(hmax, l.h);
    }
    double *V = (double *)ws.get(50, 8 * hmax * kmax);
    double *G = (double *)ws.get(51, 8 * kmax * kmax);
    double *Ystack = (double *)ws.get(52, 8 * std::max<int64_t>(ktot * t, 1));
    double *part = (double *)ws.get(53, 8 * (size_t)t * t * 8);
    size_t part_bytes = 8 * (size_t)t * t * 8;
    hipEvent_t ev[4];
    for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(ev[0], st));
    int64_t row = 0;
    for (auto &l : leaves) {
        const int64_t h = l.h, K = l.k;
        Fp64Gemm g1;  // V = Rinv Q
        g1.m = h; g1.n = K; g1.k = h;
        g1.A = l.rinv; g1.lda = h; g1.B = l.q; g1.ldb = h; g1.C = V; g1.ldc = h;
        gemm_f64(g1, st);
        Fp64Gemm g2;  // Gram = V^T V (upper)
        g2.m = K; g2.n = K; g2.k = h;
        g2.A = V; g2.lda = h; g2.ta = true; g2.B = V; g2.ldb = h; g2.C = G; g2.ldc = K; g2.tri = 1;
        gemm_f64(g2, st);
        hipLaunchKernelGGL(sampled_entries_f64, dim3((unsigned)((K * K + 255) / 256)), dim3(256), 0,
                           st, G, K, l.w, l.hbuf);
        HIP_CHECK(hipGetLastError());
        (void)row;
    }
    HIP_CHECK(hipEventRecord(ev[1], st));
    for (auto &l : leaves) {
        const int64_t K = l.k;
        potrf_f64(l.hbuf, K, K, d_info, part, part_bytes, st, rule);
        // Y rows [row, row + K) of the stack: copy B, then L^-1 B
        HIP_CHECK(hipMemcpy2DAsync(Ystack + row, 8 * ktot, l.b, 8 * K, 8 * K, t,
                                   hipMemcpyDeviceToDevice, st));
        trsm_lower_f64(l.hbuf, K, K, Ystack + row, t, ktot, st);
        row += K;
    }
    // S = C - Y^T Y (upper)
    HIP_CHECK(hipMemcpyAsync(S, C, 8 * (size_t)t * t, hipMemcpyDeviceToDevice, st));
    Fp64Gemm g3;
    g3.m = t; g3.n = t; g3.k = ktot;
    g3.A = Ystack; g3.lda = ktot; g3.ta = true; g3.B = Ystack; g3.ldb = ktot;
    g3.C = S; g3.ldc = t; g3.alpha = -1.0; g3.beta = 1.0; g3.tri = 1;
    g3.workspace = part; g3.workspace_bytes = part_bytes;
    gemm_f64(g3, st);
    HIP_CHECK(hipEventRecord(ev[2], st));
    hipLaunchKernelGGL(negate_upper_to_full, dim3((unsigned)((t * t + 255) / 256)), dim3(256), 0, st,
                       S, t, border_factor);
    HIP_CHECK(hipGetLastError());
    potrf_f64(border_factor, t, t, d_info, part, part_bytes, st, rule);
    HIP_CHECK(hipEventRecord(ev[3], st));
    HIP_CHECK(hipEventSynchronize(ev[3]));
    if (times) {
        float a = 0, b = 0, c = 0;
        HIP_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
        HIP_CHECK(hipEventElapsedTime(&c, ev[2], ev[3]));
        times->build_ms = a;
        times->leaves_ms = b;
        times->border_ms = c;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
}

}  // namespace sdpx
