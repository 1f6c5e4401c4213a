// Float64 dense kernels (see fp64_dense.hip).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdpx {

struct Fp64Gemm {
    const double *A = nullptr, *B = nullptr;
    double *C = nullptr;
    int64_t m = 0, n = 0, k = 0, lda = 1, ldb = 1, ldc = 1;
    int64_t sa = 0, sb = 0, sc = 0, batch = 1;
    double alpha = 1.0, beta = 0.0;
    bool ta = false, tb = false;
    int tri = 0;                // 0 all, 1 upper (i <= j), 2 lower (i >= j)
    double *workspace = nullptr;  // split-K partials (optional)
    size_t workspace_bytes = 0;
};

// C = alpha op(A) op(B) + beta C (column-major, batched).
void gemm_f64(const Fp64Gemm &g, hipStream_t s);
// Dynamic pivot rule of the reference's dense factor (dense_block.rs:329-365):
// diag is the unfactored diagonal on the device.
struct PivotRule {
    double eps, delta;
    const double *diag;
};
// In-place lower Cholesky; *d_info (device) = 0 or the failing column + 1.
void potrf_f64(double *A, int64_t n, int64_t lda, int *d_info, double *ws, size_t ws_bytes,
               hipStream_t s, const PivotRule *rule = nullptr);
// B <- L^-1 B and B <- L^-T B with the lower factor L.
void trsm_lower_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs, int64_t ldb,
                    hipStream_t s);
void trsm_lower_t_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs,
                      int64_t ldb, hipStream_t s);
// B <- (L L^T)^-1 B with the lower factor L.
void potrs_f64(const double *L, int64_t n, int64_t ldl, double *B, int64_t nrhs, int64_t ldb,
               double *ws, size_t ws_bytes, hipStream_t s);
// S (m x m, upper triangle, lds) = sum over svec coordinates of
// svec(Rinv A_a Rinv^T) svec(Rinv A_b Rinv^T); a_svec is s x m (s = n(n+1)/2).
void schur_psd_f64(int64_t n, int64_t m, const double *rinv, const double *a_svec, double *S,
                   int64_t lds, hipStream_t st);

// Copy the upper triangle onto the lower one.
void mirror_upper(double *A, int64_t n, int64_t lda, hipStream_t s);

// Adjoint of the PSD Schur solve (device pointers): bgrad = M^-1 ybar,
// cgrad = svec(W smat(A_svec bgrad) W), W = Rinv^T Rinv; Lm the lower factor of M.
void adjoint_psd_f64(const double *Lm, int64_t m, int64_t n, const double *rinv,
                     const double *a_svec, int64_t nrhs, const double *ybar, double *bgrad,
                     double *cgrad, hipStream_t st);

// Full adjoint at a central-path point (see fp64_dense.hip): gradients of
// L = ybar . y + <Xbar, X> + obj <C, X> with respect to b (m x nrhs), C
// (s x nrhs, svec) and, when agrad is given, every A_i (s x m x nrhs).
// ybar, xbar, obj (nrhs weights) and agrad may be null.
void adjoint_psd_full_f64(const double *Lm, int64_t m, int64_t n, const double *rinv, const double *a_svec,
                          const double *y, const double *x_svec, const double *c_svec, int64_t nrhs,
                          const double *ybar, const double *xbar, const double *obj, double *bgrad,
                          double *cgrad, double *agrad, hipStream_t st);

// One PSD block of a condensed assembly: ncols constraint columns given as
// svec columns (s x ncols) over an n x n block, its NT factor Rinv, and the
// columns' global indices (device).
struct PsdBlockF64 {
    int64_t n, ncols;
    const double *rinv, *a_svec;
    const uint64_t *columns;
};
// The condensed KKT's right-hand side xr = bx + A^T H^-1 bz and recovery
// z = H^-1 (A x - bz) over the same blocks and orthant rows as
// condensed_assemble_f64; bz and z hold the eliminated rows block after
// block (svec rows), then the orthant rows.
void condensed_rhs_f64(int64_t m, const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B, int64_t ldb,
                       const double *w, const double *bx, const double *bz, double *xr, hipStream_t st);
void condensed_recover_f64(int64_t m, const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B,
                           int64_t ldb, const double *w, const double *x, const double *bz, double *z, hipStream_t st);
void condensed_assemble_f64(int64_t m, const int64_t *p_colptr, const int64_t *p_rowval, const double *p_nzval,
                            const std::vector<PsdBlockF64> &blocks, int64_t rows, const double *B, int64_t ldb,
                            const double *w, double *S, int64_t lds, hipStream_t st);

__global__ void svec_unpack(const double *svec, int64_t s, int64_t n, int64_t m, double *out);
__global__ void svec_pack(const double *U, int64_t s, int64_t n, int64_t m, double *M);

}  // namespace sdpx
