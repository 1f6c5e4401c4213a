// GPU panel kernels of the multiple-precision LDL^T (see mp_kernels.hip).
// All pointers are device pointers to value arrays; n-limb values.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdpx {

struct LdlBlockDesc {  // mirrors the kernel's LdlBlock
    int64_t a_off, ld, n, dinv_off, first_off, sign_off, scratch_off;
};

bool mp_kernels_support(int limbs);
// Named tuning switches of the panel kernels (false: unknown name).
bool mp_set_tuning(const char *name, int64_t value);
void mp_ew_sub_packed(int limbs, void *S, int64_t ld, const void *P, int64_t n, int lower,
                      hipStream_t st);
void mp_ew_sub_rect(int limbs, void *S, int64_t lds, const void *P, int64_t ldp, int64_t m,
                    int64_t n, hipStream_t st);
void mp_ew_mul_diag(int limbs, const void *Y, int64_t ldy, const void *D, void *Z, int64_t ldz,
                    int64_t m, int64_t n, bool per_row, hipStream_t st);
void mp_sampled_entries(int limbs, const void *G, const void *w, int64_t K, void *H,
                        hipStream_t st);
// Exponent ranges (int2 min, max per row/column) of L's strictly lower
// triangle: mode 0 rows (k < j), mode 1 columns (k > j).  The row kernels
// take them (optional, may be null) as window hints; results never depend
// on them.
void mp_exp_ranges(int limbs, const void *L, int64_t ldl, int64_t n, int mode, void *out,
                   hipStream_t st);
// Y = L^-1 B of many leaves in one launch (forward_suffix_rows per leaf):
// leaf i's L (K x K, ld K) at record l_off of L_all, its envelope at
// first_off of first_all, Y (K x t, ld K, B on entry) at record y_off of
// Y_all, its row ranges (int2) at range_off of ranges_all (may be null).
struct FwdLeaf {
    int64_t l_off, first_off, k, y_off, range_off;
};
void mp_forward_rows(int limbs, const void *L_all, const int32_t *first_all, void *Y_all,
                     const void *ranges_all, const FwdLeaf *leaves, int nleaves, int64_t t, int *fail,
                     hipStream_t st);
// X L^T = B rows (n <= 128 takes a right-looking pass first and solves only
// the rows it hands back with the left-looking kernel).  work: device
// memory of at least 4 (m + 1) bytes.
void mp_trsm_base(int limbs, const void *L, int64_t ldl, int64_t n, void *B, int64_t ldb,
                  int64_t m, const int32_t *low, const void *lranges, int *fail, void *work,
                  hipStream_t st);
// status[0] = 0 ok, 1 failed pivot, 2 window overflow, 4 a bounded wait of
// the one-launch form ran out (a device error) (zero on entry);
// status[1] += regularized; status[3] = 1 + the index of a block with a
// failed pivot (status is int[4]).  A nonzero status[0] or status[2] (a
// row kernel's window overflow) on entry stops the launch.  nmax: the
// largest block order.  Each block's scratch holds ldl_scratch_records(n)
// records.
void mp_ldl_dots(int limbs, void *A, void *dinv, const int32_t *first, const int8_t *signs,
                 int sign_mode, const void *blocks, int nblocks, int64_t nmax, void *scratch,
                 const void *reg_eps, const void *reg_delta, int *status, hipStream_t st);
// r and w triangles, the pivots, per-row exponent ranges (16 bytes a row,
// within one record of any width), two per-row exact-dot windows (slot
// totals and base, 8 (2W + 21) bytes a row each: within 10 records of any
// width), and two int flags a row (the one-launch form's order)
inline size_t ldl_scratch_records(size_t n) { return n * (n + 1) + 2 * n + 20 * n + n; }
void mp_trsm_base_n(int limbs, const void *L, int64_t ldl, int64_t n, void *B, int64_t ldb,
                    int64_t m, const void *lranges, int *fail, hipStream_t st);
void mp_transpose(int limbs, const void *in, int64_t ldi, int64_t m, int64_t n, void *out,
                  int64_t ldo, hipStream_t st);
void mp_envelope(int limbs, const void *A, int64_t ld, int64_t n, int32_t *first, hipStream_t st);
void mp_mirror_upper(int limbs, void *S, int64_t n, int64_t ld, hipStream_t st);
// the strictly lower triangle of an n x n block copied onto its upper one
void mp_mirror_lower(int limbs, void *S, int64_t n, int64_t ld, hipStream_t st);
// out[c] (ncols records) = max_i |X[i + c ldx]| over m rows, +infinity when
// the column holds a NaN or an infinity (norm_inf of each column)
void mp_col_absmax(int limbs, const void *X, int64_t ldx, int64_t m, int64_t ncols, void *out, hipStream_t st);
// FMA sweeps of the arrow solves (DenseLeaf forward_many / backward order):
// descs (device) lists ndesc factors, each an order-n unit lower L (ld n) at
// record l_off of L_all and right-hand sides x (column c at record x_off +
// c ldx of X), one workgroup per (factor, column).  Forward: the sweep,
// then w stored to Wout and RN(w D^-1) (D at record d_off of dinv_all) to
// Vout (column c at x_off + c ldw), or in place (Wout null) when Vout is
// null.  Backward: blocks of `back` rows from the bottom.  Orders up to 8192.
struct SweepDescHost {
    int64_t l_off, n, x_off, ldx, d_off;
};
bool mp_sweep_supports(int64_t n);
void mp_arrow_forward(int limbs, const void *L_all, const void *dinv_all, const void *descs, int ndesc,
                      int64_t nmax, int64_t cols, void *X, void *Wout, void *Vout, int64_t ldw, hipStream_t st);
void mp_arrow_backward(int limbs, const void *L_all, const void *descs, int ndesc, int64_t nmax, int64_t cols,
                       void *X, int back, hipStream_t st);
// The same sweeps for one factor L (n x n, ld n) of any order, its
// right-hand sides X (n x cols, ldx) in place: blocks of 32 rows forward,
// of `back` (<= 32) rows from the bottom backward, each block's own rows
// on one wave and every other row's updates spread over the chip.
void mp_sweep_blocked(int limbs, const void *L, int64_t n, int64_t cols, void *X, int64_t ldx, bool backward,
                      int back, hipStream_t st);
void mp_zero_values(int limbs, void *P, int64_t count, hipStream_t st);
// hi = RN(d x), lo = d x - hi (exact) for an m x n block of X; d per row
// (mode 0), per column (1) or a single value (2)
void mp_two_product(int limbs, const void *X, int64_t ldx, const void *D, int mode, void *Hi, void *Lo, int64_t ldo,
                    int64_t m, int64_t n, hipStream_t st);
// svec_quadratic's weights from an svec x and the constant s (device):
// Hi + Lo = diag(x_ii) + s triu(x) exactly (h x h each, zero below)
void mp_quadratic_weights(int limbs, const void *P, int64_t h, const void *S, void *Hi, void *Lo, hipStream_t st);
// the full symmetric h x h matrix of a column-packed upper triangle
void mp_unpack_symmetric(int limbs, const void *P, int64_t h, void *out, hipStream_t st);
void mp_abs_values(int limbs, void *X, int64_t count, hipStream_t st);
// C = axpby(alpha, P, beta, C) as the reference's MPFR axpby; codes 0 zero,
// 1 one, 2 minus one, 3 other (alpha, beta: one device value each, read only
// for code 3); tri 0 all, 1 upper, 2 lower; packed: C is column-packed upper;
// sym: P holds a symmetric matrix's upper triangle
void mp_axpby(int limbs, const void *P, int64_t ldp, int64_t m, int64_t n, int acode, const void *alpha, int bcode,
              const void *beta, void *C, int64_t ldc, int tri, bool packed, bool sym, hipStream_t st);
// One block of the grouped exact product (device pointers; cols ascending
// global columns, int64)
struct BlockScatter {
    const void *y, *d;
    int64_t rows, ncols, row0;
    const int64_t *cols;
};
// Y, Z (K x m, ld K, zero beforehand): every block's y and RN(diag(d) y) at
// its rows and global columns (descs on the device)
void mp_block_scatter(int limbs, const BlockScatter *descs, int nblocks, int64_t max_elems, void *Z, void *Y,
                      int64_t K, hipStream_t st);
// out[p] = G[ia[p] + ib[p] ldg] (ia, ib device)
void mp_gather(int limbs, const void *G, int64_t ldg, const int64_t *ia, const int64_t *ib, int64_t count, void *out,
               hipStream_t st);
// X = RN(RN(Wm - S) D) rows i (D per row) of an m x n block
void mp_arrow_second(int limbs, const void *Wm, int64_t ldw, const void *S, int64_t lds, const void *D, void *X,
                     int64_t ldx, int64_t m, int64_t n, hipStream_t st);
// Exact-dot Cholesky (xpotrf), one launch per column: status[0] = info
// (failing column + 1), status[2] |= 1 when a dot's exponent spread exceeds
// the window; both must be zero on entry.
void mp_potrf(int limbs, void *A, int64_t n, int64_t lda, int lower, int *status, hipStream_t st);
// xtrsm_lower (a: n x n, ld n; b: n x nrhs, ld n); fail |= window overflow.
void mp_trsm_lower(int limbs, const void *A, int64_t n, void *B, int64_t nrhs, int transpose, int *fail,
                   hipStream_t st);
// xpotrs with a factor from xpotrf (one lane per right-hand side).
void mp_potrs(int limbs, const void *A, int64_t n, int64_t lda, int lower, void *B, int64_t ldb,
              int64_t nrhs, hipStream_t st);

}  // namespace sdpx
