/* sdpx_gpu: MI355X (gfx950) backend of SDPX's per-iteration Schur-complement
 * hot path, behind a C ABI a thin Rust shim binds from the solver's traits.
 *
 * Conventions follow include/sdpx.h:8-17 and crates/ffi/src/lib.rs:30-43 of
 * the reference: integer return codes, a thread-local last error, column-major
 * storage, transposes given as the bytes 'N' / 'T' ('C' is read as 'T').
 *
 * Return codes: 0 OK, 1 invalid input, 2 declined (the exact residue product
 * does not admit a plan: non-finite entry, operand exponent spread above 4096
 * bits, or integer widths beyond the kernels' limits; outputs are unwritten
 * and the caller falls back exactly as on a `false` from the reference's
 * exact hooks), 3 busy, 5 device or internal error, 6 insufficient output
 * capacity.
 *
 * Multiple-precision values use the interchange layout of
 * MpFloat<N>::exact_encode (crates/arithmetic/src/lib.rs:509-533): per value
 * N little-endian u64 limbs, then i32 kind (MPFR kind code: 0 NaN, 1 Inf,
 * 2 zero, 3 regular, negated for a negative sign), then i32 exponent; a
 * regular value is M * 2^(exponent - 64N) with 2^(64N-1) <= M < 2^(64N).
 * Arrays are 8N+8 bytes per value, with no padding.  1 <= N <= 16.
 *
 * Host-pointer entry points stage through device memory; the `_dev` variants
 * take device pointers and run on the library's stream (returned by
 * sdpx_gpu_stream) and return once the results are complete.  That stream
 * is created non-blocking, so it does not wait for the legacy default
 * stream: a caller that writes a `_dev` input on another stream must make
 * that write complete (a stream synchronize, or an event the library
 * stream waits on) before the call.
 */
#ifndef SDPX_GPU_H
#define SDPX_GPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SDPX_GPU_ABI_VERSION 5u
#define SDPX_GPU_OK 0
#define SDPX_GPU_INVALID 1
#define SDPX_GPU_DECLINED 2
#define SDPX_GPU_BUSY 3
#define SDPX_GPU_DEVICE_ERROR 5
#define SDPX_GPU_CAPACITY 6

/* ABI version of this header (no reference counterpart; mirrors
 * SDPX_ABI_VERSION, include/sdpx.h:19). */
uint32_t sdpx_gpu_abi_version(void);

/* Select and initialise the HIP device (default 0).  Returns 5 when no
 * device is present. */
int32_t sdpx_gpu_init(int32_t device);

/* The library's HIP stream (hipStream_t) for the `_dev` entry points. */
void *sdpx_gpu_stream(void);

/* Thread-local most recent error, as sdpx_last_error (include/sdpx.h:111-113):
 * copies a NUL-terminated message, returns the bytes needed including NUL. */
uint64_t sdpx_gpu_last_error(char *out, uint64_t capacity);

/* Statistics of the most recent exact product on this thread: moduli used,
 * INT8 planes per operand, and log2 of the modulus product.  For
 * measurement; no reference counterpart. */
int32_t sdpx_gpu_last_plan(uint32_t *moduli, uint32_t *planes, double *bits);

/* Measurement only (no reference counterpart): when on, later exact
 * products record HIP events around their stages; sdpx_gpu_last_stage_times
 * returns this thread's most recent stage times in ms: scan, encode,
 * product (all INT8 MFMA launches), crt. */
int32_t sdpx_gpu_set_profiling(int32_t on);

/* Named tuning switches of the GPU panel kernels, for measurement and tests;
 * no reference counterpart.  "trsm_rl" (0/1): the right-looking pass of the
 * triangular-solve base (default 1); "trsm_rl_room" (bits, default 128): its
 * exact windows' room below their first term, a negative value sends every
 * row to the left-looking kernel; "ldl_one_launch" (0/1/2): ldl_dots as
 * one launch of row workgroups for a single block (1, default), for every
 * call with all blocks' rows interleaved (2), or one launch per column
 * (0); "sweep_wave" (0/1): the arrow solve's single-rounding updates on
 * whole waves (default 1) or one lane a row; "fwd_group" (4, 8, 16, 32):
 * lanes per exact dot of the forward rows (default 8).  Results never
 * depend on any of them.  Returns 1 for an unknown name or value. */
int32_t sdpx_gpu_set_tuning(const char *name, int64_t value);
int32_t sdpx_gpu_last_stage_times(double *ms, uint32_t count);
/* Totals over this thread's profiled exact products since the last reset:
 * stage ms (4, as above), algorithmic INT8 operations (2 x outputs x k x
 * INT8 GEMMs per exact product), multiple-precision MACs, calls; reset != 0
 * clears them after reading. */
int32_t sdpx_gpu_rns_totals(int32_t reset, double *stage_ms4, double *int8_ops, double *mp_macs,
                            int64_t *calls);

/* ---- Exact residue (RNS) products on INT8 matrix cores ----------------
 *
 * sdpx_gpu_rns_gemm replaces rns_blas::gemm
 * (crates/solver/src/algebra/dense/blas/rns_blas.rs:2134-2417), the kernel
 * behind XgemmScalar::xgemm_upper_exact and the exact branch of xgemm/xsyrk
 * (blas/traits.rs:228-249, blas/mpfr.rs:245-348, 540-590):
 *   out = op(a) * op(b), m x n, inner k; every entry the exact dot product
 *   rounded once to nearest-even at 64N bits.
 * Output layout as the reference: when upper_only, m == n and
 * out_len == n(n+1)/2 the upper triangle is column-packed
 * (entry (i, j), i <= j, at j(j+1)/2 + i); otherwise out has ldc = m and
 * holds m*n values, of which only i <= j are written when upper_only.
 * An operand that is entirely zero gives zeros.  Returns 2 (declined, out
 * untouched) exactly where the reference's kernel returns false for
 * non-finite entries or spreads above 4096 bits, and additionally when an
 * operand's integer width exceeds the GPU kernels' limit (1536 bits). */
int32_t sdpx_gpu_rns_gemm(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m, uint64_t n,
                          uint64_t k, const void *a, uint64_t lda, const void *b, uint64_t ldb,
                          uint32_t upper_only, void *out, uint64_t out_len);
/* Same on device pointers. */
int32_t sdpx_gpu_rns_gemm_dev(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m,
                              uint64_t n, uint64_t k, const void *a, uint64_t lda, const void *b,
                              uint64_t ldb, uint32_t upper_only, void *out, uint64_t out_len);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sdpx_gpu_rns_gemm_dev with every temporary (operand
 * scans, INT8 residue planes, the CRT's residue buffer, plan tables) in one
 * fixed array of SDPX_GPU_NOALLOC_SCRATCH_BYTES per device, compiled into
 * the library and held until the process ends, so the call allocates
 * nothing (no hipMalloc, hipFree or pinned host memory).  Inputs and outputs
 * are the caller's device buffers.  Returns 6 (capacity; output untouched,
 * with a message) when the product's temporaries exceed the array: about
 * G (m + n) k + 2 L m n bytes (dimensions padded to 128 or 256, k to 64)
 * for L moduli and G INT8 planes, roughly L = 65, G = 100 for full 256-bit
 * operands and L = 120, G = 150 at 512 bits, so 256 x 256 x 256 fits at
 * either width.  Every call on a device
 * shares that device's array and leaves its contents behind; the library's
 * device lock serialises calls within a process, so only calls that
 * overlap on one device from outside that lock would race. */
#define SDPX_GPU_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t sdpx_gpu_rns_gemm_dev_noalloc(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m,
                                      uint64_t n, uint64_t k, const void *a, uint64_t lda,
                                      const void *b, uint64_t ldb, uint32_t upper_only, void *out,
                                      uint64_t out_len);

/* Grouped exact blocks: replaces rns_blas::gemm_blocks_upper
 * (rns_blas.rs:2671-2922) behind XgemmScalar::xgemm_blocks_upper_exact
 * (traits.rs:250-258): c[i + j*m] for i <= j receives the upper triangle of
 * sum_b Z_b^T Y_b over blocks of uneven height on sorted global columns,
 * Z_b = RN(diag(d_b) Y_b) entrywise, each entry the exact sum over all
 * blocks rounded once.  y is column-major rows x ncols.  Returns 2 on a
 * decline (outputs may be partially written, as the reference allows). */
typedef struct {
    uint64_t struct_size;       /* sizeof(sdpx_gpu_exact_block) */
    uint64_t rows, ncols;
    const void *y;              /* rows x ncols values */
    const void *d;              /* rows values */
    const uint64_t *columns;    /* ncols ascending global columns < m */
} sdpx_gpu_exact_block;
int32_t sdpx_gpu_rns_gemm_blocks_upper(uint32_t limbs, uint64_t m, uint64_t nblocks,
                                       const sdpx_gpu_exact_block *blocks, void *out);
/* Same with each block's y and d and the output on the device (columns stay
 * host arrays); 128, 256 and 512 bits.  The blocks are scattered into HBM
 * and summed by one exact product launch. */
int32_t sdpx_gpu_rns_gemm_blocks_upper_dev(uint32_t limbs, uint64_t m, uint64_t nblocks,
                                           const sdpx_gpu_exact_block *blocks, void *out);

/* General products with scaling, replacing the MPFR XgemmScalar::xgemm and
 * XsyrkScalar::xsyrk (blas/mpfr.rs:245-348 gemm, 540-626 syrk):
 *   xgemm: c (ldc) = axpby(alpha, RN(op(a) op(b)), beta, c)
 *   xsyrk: the uplo triangle of c = axpby(alpha, RN(op(a) op(a)^T), beta, c)
 *          (op(a) n x k; c column-packed upper when uplo = 'U', ldc = n and
 *          c_len = n(n+1)/2, else ldc >= n and the other triangle untouched)
 * with the reference's axpby (mpfr.rs:61-80): p = 0, x, -x or RN(alpha x)
 * for alpha 0, 1, -1 or other; then p, RN(p + c), RN(p - c) or
 * RN(p + RN(beta c)) for beta 0, 1, -1 or other.  alpha = 0 computes no
 * product.  alpha and beta are one host value each, also in the _dev forms.
 * Returns 2 (declined, c untouched) where the exact product declines. */
int32_t sdpx_gpu_xgemm(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m, uint64_t n, uint64_t k,
                       const void *alpha, const void *a, uint64_t lda, const void *b, uint64_t ldb, const void *beta,
                       void *c, uint64_t ldc);
int32_t sdpx_gpu_xgemm_dev(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m, uint64_t n, uint64_t k,
                           const void *alpha, const void *a, uint64_t lda, const void *b, uint64_t ldb,
                           const void *beta, void *c, uint64_t ldc);
int32_t sdpx_gpu_xsyrk(uint32_t limbs, uint8_t uplo, uint8_t trans, uint64_t n, uint64_t k, const void *alpha,
                       const void *a, uint64_t lda, const void *beta, void *c, uint64_t ldc, uint64_t c_len);
int32_t sdpx_gpu_xsyrk_dev(uint32_t limbs, uint8_t uplo, uint8_t trans, uint64_t n, uint64_t k, const void *alpha,
                           const void *a, uint64_t lda, const void *beta, void *c, uint64_t ldc, uint64_t c_len);

/* ---- Exact multilinear forms -------------------------------------------
 *
 * The residue-BLAS quadratic forms behind XgemmScalar (blas/traits.rs:
 * 259-296, blas/mpfr.rs:410-430): every output entry is the exact value of
 * the form rounded once to nearest-even at 64N bits (128, 256 or 512 bits
 * here).  Column-major operands; 2 = declined (the caller takes its own
 * path; outputs may be partially written, as the reference allows).  The
 * _dev forms take device pointers and run on the library stream. */

/* diag_congruence (rns_blas.rs:2997-3195), xdiag_congruence_upper_exact:
 * c[i + j*m] = (a^T diag(d) a)_ij for i <= j, a k x m (ld k), d k values;
 * the strict lower triangle is left untouched. */
int32_t sdpx_gpu_diag_congruence_upper_exact(uint32_t limbs, uint64_t m, uint64_t k, const void *a,
                                             const void *d, void *c);
/* Same on device pointers; upper_only = 0 also writes the lower triangle. */
int32_t sdpx_gpu_diag_congruence_upper_exact_dev(uint32_t limbs, uint64_t m, uint64_t k, const void *a,
                                                 const void *d, void *c, uint32_t upper_only);
/* congruence (rns_blas.rs:3563-3790), xcongruence_exact: c (ldc = m) =
 * op(a) x op(a)^T for x k x k (ld ldx), op(a) m x k; upper triangle only
 * when upper_only (then the strict lower triangle is left untouched). */
int32_t sdpx_gpu_congruence_exact(uint32_t limbs, uint8_t transa, uint64_t m, uint64_t k, const void *a,
                                  uint64_t lda, const void *x, uint64_t ldx, void *c, uint32_t upper_only);
int32_t sdpx_gpu_congruence_exact_dev(uint32_t limbs, uint8_t transa, uint64_t m, uint64_t k, const void *a,
                                      uint64_t lda, const void *x, uint64_t ldx, void *c, uint32_t upper_only);
/* svec_quadratic (rns_blas.rs:4032-4173), xsvec_quadratic_exact:
 * out[k] = sum_{i <= j} c_ij x_t q_ik q_jk for k < kmax, x an svec (entry
 * (i, j), i <= j, at t = j(j+1)/2 + i), c_ii = 1 and c_ij = *sqrt2 (one
 * value) otherwise; q h x kmax (ld h), read as |q| when abs_q. */
int32_t sdpx_gpu_svec_quadratic_exact(uint32_t limbs, uint64_t h, uint64_t kmax, const void *q, uint32_t abs_q,
                                      const void *x, const void *sqrt2, void *out);
int32_t sdpx_gpu_svec_quadratic_exact_dev(uint32_t limbs, uint64_t h, uint64_t kmax, const void *q,
                                          uint32_t abs_q, const void *x, const void *sqrt2, void *out);
/* symmetric_bilinear (rns_blas.rs:4175-4276), xsymmetric_bilinear_exact:
 * out[p] = q_a^T X q_b for (a, b) = (pairs[2p], pairs[2p+1]) (host array,
 * indices < columns, else 1), X symmetric h x h given as its column-packed
 * upper triangle (no svec factors), q h x columns (ld h).  Declined for
 * h = 0, as the reference's form. */
int32_t sdpx_gpu_symmetric_bilinear_exact(uint32_t limbs, uint64_t h, uint64_t columns, const void *q,
                                          const void *x, uint64_t npairs, const uint64_t *pairs, void *out);
int32_t sdpx_gpu_symmetric_bilinear_exact_dev(uint32_t limbs, uint64_t h, uint64_t columns, const void *q,
                                              const void *x, uint64_t npairs, const uint64_t *pairs, void *out);

/* ---- Multiple-precision dense LDL^T -----------------------------------
 *
 * sdpx_gpu_mp_ldl replaces DenseLeaf::factor_leaf / factor_loaded of the
 * arrow backend (crates/solver/src/solver/kkt/ldl/arrow.rs:808-926): from
 * order 256 the recursive factor ldl_rec (255-415) with its panel solve
 * solve_unit_lower_t (417-490), below it ldl_dots (156-253), bitwise as the
 * reference.  The lower triangle of a (n x n, lda, values of 64*limbs bits)
 * is overwritten by the unit lower factor (ones on the diagonal), dinv
 * (n values) receives D^-1.  signs: 0 entries (all +1), 1 (shared) or n
 * (per pivot, +1/-1); reg_eps / reg_delta (one value each, or both NULL)
 * select the pivot regularization of DenseLeaf::pivot (618-634).  *ok = 0
 * on a failed pivot (the reference's Err; contents unspecified).  As
 * factor_loaded (arrow.rs:819-846), a recursive factor that fails on a pivot
 * (one rounding per level can cancel a pivot to an exact zero that exact
 * dots keep nonzero) is redone from the input by ldl_dots, and only that
 * factor's regularization count is kept.  Every
 * exact product of the recursion runs as an RNS product on the GPU.  flags:
 * SDPX_GPU_LDL_DOTS forces ldl_dots at any order (the reference's
 * factor_dots fallback); SDPX_GPU_LDL_HOST_PRODUCTS keeps every product on
 * host exact dots (same values; for testing). */
#define SDPX_GPU_LDL_DOTS 1u
#define SDPX_GPU_LDL_HOST_PRODUCTS 2u
int32_t sdpx_gpu_mp_ldl(uint32_t limbs, uint64_t n, void *a, uint64_t lda, const int8_t *signs,
                        uint64_t nsigns, const void *reg_eps, const void *reg_delta, void *dinv,
                        uint32_t flags, int32_t *ok);
/* Device-pointer form of sdpx_gpu_mp_ldl (128/256/512 bits; signs and the
 * regularization values stay host pointers).  Returns 2 (declined) when a
 * dot's exponent spread or a product exceeds the GPU limits; a declined
 * call leaves a and dinv unspecified (the caller refactors from its own
 * copy of the input). */
int32_t sdpx_gpu_mp_ldl_dev(uint32_t limbs, uint64_t n, void *a, uint64_t lda, const int8_t *signs,
                            uint64_t nsigns, const void *reg_eps, const void *reg_delta, void *dinv,
                            uint32_t flags, int32_t *ok);
/* L D L^T x = b for nrhs columns (b: n x nrhs, ldb, overwritten by x) with
 * the factor of sdpx_gpu_mp_ldl (unit lower l, dinv).  Replaces
 * DenseLeaf::solve_many (arrow.rs:1155-1195) for a dense leaf; the rounding
 * sequence is the recursive panel form (exact dots below 128 columns, one
 * exact product per level, as forward_panel_rows, arrow.rs:1020-1100), not
 * the reference's per-entry FMA chains, so results agree to the working
 * precision rather than bitwise.  128/256/512 bits; returns 2 when a dot's
 * exponent spread exceeds the GPU window. */
int32_t sdpx_gpu_mp_ldl_solve(uint32_t limbs, uint64_t n, const void *l, uint64_t ldl,
                              const void *dinv, uint64_t nrhs, void *b, uint64_t ldb);
/* Device-pointer form of sdpx_gpu_mp_ldl_solve (l, dinv, b in HBM). */
int32_t sdpx_gpu_mp_ldl_solve_dev(uint32_t limbs, uint64_t n, const void *l, uint64_t ldl,
                                  const void *dinv, uint64_t nrhs, void *b, uint64_t ldb);
/* Exact-dot Cholesky of the stored triangle (uplo 'L': a[i + k lda], i >= k;
 * 'U': a[k + i lda]) in place.  Replaces XpotrfScalar::xpotrf for MPFR
 * (crates/solver/src/algebra/dense/blas/mpfr.rs:727-789): each pivot is
 * d = RN(a_jj - RN(sum_k L_jk^2)), stored, then rejected unless finite and
 * positive (info = j + 1, the rest untouched), else replaced by RN(sqrt d);
 * each entry is RN(RN(a_ij - RN(sum_k L_ik L_jk)) / d).  info -1: bad uplo,
 * -4: lda below n.  128/256/512 bits; returns 2 (nothing written) for
 * non-finite input or a dot whose exponent spread exceeds the GPU window. */
int32_t sdpx_gpu_mp_potrf(uint32_t limbs, uint8_t uplo, uint64_t n, void *a, uint64_t lda, int32_t *info);
/* Solve A X = B with the factor of sdpx_gpu_mp_potrf.  Replaces
 * XpotrsScalar::xpotrs (mpfr.rs:790-872): per right-hand side the forward
 * and backward chains v = RN(v - a_ik x_k) (one rounding, the fused
 * multiply-add) and x_i = RN(v / a_ii), in the reference's order.  info -1:
 * bad uplo, -5: lda below n, -7: ldb below n.  Returns 2 for non-finite
 * input. */
int32_t sdpx_gpu_mp_potrs(uint32_t limbs, uint8_t uplo, uint64_t n, uint64_t nrhs, const void *a,
                          uint64_t lda, void *b, uint64_t ldb, int32_t *info);
/* L X = B (or L^T X = B) with the lower triangle of the n x n array a (ld n)
 * and B n x nrhs (ld n).  Replaces the default xtrsm_lower for MPFR
 * (crates/solver/src/algebra/dense/blas/traits.rs:85-113): x_i =
 * RN(RN(b_i - RN(exact dot of the solved entries)) / L_ii).  Returns 2
 * (nothing written) for non-finite input or a dot beyond the window. */
int32_t sdpx_gpu_mp_trsm_lower(uint32_t limbs, uint64_t n, const void *a, void *b, uint64_t nrhs,
                               uint32_t transpose);
/* Counters of this thread's last sdpx_gpu_mp_ldl / _dev (measurement and
 * tests; each pointer optional): exact products on the GPU and on host
 * dots, regularized pivots, GPU multiple-precision MACs, declined (1 when
 * the device panels declined and the host panels produced the result), and
 * retried (1 when the recursive factor failed and exact dots redid it). */
int32_t sdpx_gpu_mp_ldl_stats(int64_t *gpu_products, int64_t *host_products, int64_t *regularized,
                              double *gpu_mp_macs, int64_t *declined, int64_t *retried);

/* ---- Sampled Schur build + arrow factor (multiple precision) ------------
 *
 * One IPM iteration's Schur build and factor of a fully sampled condensed
 * KKT (factor-authoritative dim-1 sampled PSD blocks, sdpx_sampled_block of
 * include/sdpx.h, coupled to t border rows), bitwise as the reference:
 *   V = Rinv Q, Gram = V^T V            (sampled/mod.rs:1836-1868, 1990-2143)
 *   H_pq = w_p w_q 1/2 (K_pq^2 + K_pq^2) (entry, sampled/mod.rs:1973-1988)
 *   H = L D L^T, Y = L^-1 B             (Leaf::refactor, arrow.rs:1522-1640;
 *                                        ldl_dots 156-253, forward_suffix_rows
 *                                        985-1019)
 *   S -= Y^T RN(D^-1 Y) per leaf in order (contribution_product /
 *                                        subtract_contributions, 1935-2074)
 *   S = L_b D_b L_b^T                   (factor_border 3378-3404, ldl_rec)
 * Leaf arrays are column-major values of 64*limbs bits.  border (t x t,
 * ld t) holds the border block C in its upper triangle on entry and the
 * border factor in its lower triangle on return; border_dinv receives D_b^-1.
 * factors (optional, one per leaf) receive L (k x k), D^-1 (k), Y (k x t).
 * Every exact product runs as an RNS product on the GPU.  Leaves of 256 or
 * more samples are declined (2, nothing written): the reference factors them
 * recursively and solves their Y by recursive panels (factor_loaded,
 * arrow.rs:819; Leaf::refactor 1596-1632), which this pipeline does not.
 * The border factor retries by exact dots as sdpx_gpu_mp_ldl does. */
typedef struct {
    uint64_t struct_size;        /* sizeof(sdpx_gpu_sampled_leaf) */
    uint64_t side, samples;      /* h (basis rows), k (samples) */
    const void *rinv;            /* h x h */
    const void *basis;           /* h x k */
    const void *weights;         /* k */
    const void *coupling;        /* k x t */
} sdpx_gpu_sampled_leaf;
typedef struct {
    uint64_t struct_size;        /* sizeof(sdpx_gpu_sampled_factor) */
    void *l, *dinv, *y;          /* outputs, each optional */
} sdpx_gpu_sampled_factor;
int32_t sdpx_gpu_sampled_arrow(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                               uint64_t t, void *border, const int8_t *border_signs,
                               uint64_t nborder_signs, void *border_dinv,
                               sdpx_gpu_sampled_factor *factors, uint32_t flags, int32_t *ok);
/* Device-resident form (all pointers device pointers; 128/256/512 bits):
 * leaf factors, D^-1 and Y land in l_all (sum k^2 values), dinv_all (sum k)
 * and y_all (sum k*t) in leaf order.  Returns 2 (declined, nothing
 * guaranteed) when a dot's exponent spread or a product exceeds the GPU
 * kernels' limits; the host entry point above then finishes on the host. */
int32_t sdpx_gpu_sampled_arrow_dev(uint32_t limbs, uint64_t nleaves,
                                   const sdpx_gpu_sampled_leaf *leaves, uint64_t t, void *l_all,
                                   void *dinv_all, void *y_all, void *border,
                                   const int8_t *border_signs, uint64_t nborder_signs,
                                   void *border_dinv, int32_t *ok);
/* Measurement only (no reference counterpart): with profiling on
 * (sdpx_gpu_set_profiling), the device LDL^T / sampled pipelines of this
 * thread time their launches by kernel category; returns the sums in ms
 * (and the timed launch groups) since the last reset: 0 ldl_dots panels
 * (the pivot chains), 1 triangular-solve bases, 2 forward rows (Y = L^-1 B),
 * 3 exact products (all RNS stages).  reset != 0 clears them. */
int32_t sdpx_gpu_panel_phase_times(double *ms4, int64_t *launches4, int32_t reset);

/* ArrowLDLSolver::solve_many (crates/solver/src/solver/kkt/ldl/arrow.rs:
 * 3793-3980) on the factors of sdpx_gpu_sampled_arrow_dev (device pointers
 * except samples, a host array): samples[i] is leaf i's order K_i; l_all, dinv_all,
 * y_all, border (its lower triangle the unit factor) and border_dinv as
 * that call left them.  x (n x nrhs, ldx >= n, n = sum K_i + t) holds the
 * right-hand sides on entry (each column: leaf 0's K_0 rows, ..., leaf
 * last's, then the t border rows) and the solutions on return.  Bitwise
 * the reference's sequence for leaves whose every row couples: per leaf
 * w = L^-1 b by single-rounding FMA updates in DenseLeaf::forward_many's
 * order (1155-1164), v = RN(w D^-1) (first_many 1682-1712); each border
 * entry minus RN(Y_l^T v_l), one exact dot per leaf, leaf by leaf (the
 * coupling of 3857-3898); the border factor's forward sweep, D_b^-1 and
 * backward sweep in blocks of 32 (DenseLeaf::solve_many, backward 1165-1183,
 * FORWARD_BLOCK); per leaf RN(RN(w - RN(Y_l y)) D^-1) and the backward
 * sweep in blocks of 16 (second_many 1713-1745, LEAF_BLOCK).  Orders up
 * to 8192; returns 2 (declined, x unspecified) when a product declines. */
int32_t sdpx_gpu_sampled_arrow_solve_dev(uint32_t limbs, uint64_t nleaves, const uint64_t *samples, uint64_t t,
                                         const void *l_all, const void *dinv_all, const void *y_all,
                                         const void *border, const void *border_dinv, uint64_t nrhs, void *x,
                                         uint64_t ldx);

/* The sampled Schur build alone, device pointers: SampledSchurWorkspace::
 * update_with_pool and entry (crates/solver/src/solver/sampled/mod.rs:
 * 1836-2143, 1973-1988) for every leaf, leaf i's H (K_i x K_i, lower
 * triangle, zeros above) at record sum_{j<i} K_j^2 of h_all: the values
 * sdpx_gpu_sampled_arrow_dev factors.  Only side, samples, rinv, basis and
 * weights of each leaf are read.  2: a product declined. */
int32_t sdpx_gpu_sampled_schur_dev(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                                   void *h_all);

/* DirectLDLKKTSolver::solve_many with stationary iterative refinement
 * (crates/solver/src/solver/kkt/direct/solver.rs:470-520; refine 1046-1129;
 * settings of default/settings.rs:252-270) on the sampled arrow KKT
 * K = [blockdiag(H_i) B; B^T C], device pointers except settings, ok, norms
 * and refinements (host): x = K^-1 b by sdpx_gpu_sampled_arrow_solve_dev's
 * sweeps, then per column, while its residual norm exceeds
 * abstol + reltol |b|_inf and at most max_iter times: dx = K^-1 e,
 * x' = RN(x + dx), e' = b - K x'; x' is kept unless |e|/|e'| <= 1, and a
 * ratio below stop_ratio ends the column.  Every residual row is
 * RN(b_i - sum_j K_ij x_j), exact and rounded once (ExactRows::residual,
 * 1486-1670), computed as one exact product [K | b] [-x; e_c] on the INT8
 * residue engine.  h_all from sdpx_gpu_sampled_schur_dev; leaves[i].coupling
 * is leaf i's B (K_i x t); c holds the border block C in its upper triangle
 * (t x t, ld t; the input of sdpx_gpu_sampled_arrow_dev, not its output);
 * l_all .. border_dinv are the factors sdpx_gpu_sampled_arrow_dev left.
 * b (n x nrhs, ldb >= n, n = sum K_i + t) is read, x (ldx >= n) written;
 * ok[c] is 1 when column c stayed finite, norms (nrhs records) receive the
 * infinity norm of the returned column's residual, refinements the
 * correction solves over all columns.  Workspace: (n + nrhs) n records for
 * [K | b].  2: a product declined (x unspecified). */
typedef struct {
    uint64_t struct_size;        /* sizeof(sdpx_gpu_refine_settings) */
    const void *reltol, *abstol; /* records of the solve's precision */
    const void *stop_ratio;      /* record */
    uint32_t max_iter;
} sdpx_gpu_refine_settings;
int32_t sdpx_gpu_sampled_kkt_solve_dev(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                                       uint64_t t, const void *h_all, const void *c, const void *l_all,
                                       const void *dinv_all, const void *y_all, const void *border,
                                       const void *border_dinv, const sdpx_gpu_refine_settings *settings,
                                       uint64_t nrhs, const void *b, uint64_t ldb, void *x, uint64_t ldx,
                                       int32_t *ok, void *norms, int64_t *refinements);

/* Phase times (s: build, leaves, border), product counts, declined (1 when
 * the device pipeline declined and host panels produced the result),
 * retried (border factors redone by exact dots) and the index of a leaf
 * whose factor failed (-1: none) of this thread's last
 * sdpx_gpu_sampled_arrow / _dev (measurement and tests; pointers optional). */
int32_t sdpx_gpu_sampled_stats(double *seconds3, int64_t *gpu_products, int64_t *host_products,
                               int64_t *declined, int64_t *retried, int64_t *failed_leaf);

/* Float64 form of the same Schur build + arrow factor on device pointers
 * (the binary64 dense elimination H = L L^T, Y = L^-1 B, S = C + Y^T Y of
 * kkt/ldl/dense_block.rs:95-229, with the leaves H_pq = w_p w_q K_pq^2):
 * s (t x t, upper) = c - sum Y^T Y, border_factor (t x t) = chol(-s) lower.
 * dyn_delta > 0 applies the dynamic pivot rule (pivots <= dyn_eps become
 * dyn_delta, as the reference's dynamic regularization) to every factor.
 * info: 0, or the failing pivot column + 1 of the last failing factor.
 * times_ms (optional, 3): build, leaves, border. */
typedef struct {
    uint64_t struct_size;        /* sizeof(sdpx_gpu_sampled_leaf_f64) */
    uint64_t side, samples;
    const double *rinv, *basis, *weights, *coupling;  /* device */
} sdpx_gpu_sampled_leaf_f64;
int32_t sdpx_gpu_sampled_arrow_f64_dev(uint64_t nleaves, const sdpx_gpu_sampled_leaf_f64 *leaves,
                                       uint64_t t, const double *c, double *s, double *border_factor,
                                       double dyn_eps, double dyn_delta, int32_t *info,
                                       double *times_ms);

/* ---- Float64 dense factor / solve -------------------------------------
 *
 * A factor handle replaces the dense Float64 backend behind
 * DirectLDLSolver (crates/solver/src/solver/kkt/direct/mod.rs:11-62),
 * DenseBlockSolver::refactor / solve (kkt/ldl/dense_block.rs:784-851,
 * tiled_potrf 515-594, tiled_trsv 604-669): refactor reads the upper
 * triangle of a symmetric n x n matrix (column-major, lda) and factors it
 * as L L^T on the GPU; info = 0 on success, else the 1-based column of the
 * first non-positive pivot (LAPACK dpotrf convention; the caller falls back
 * or escalates regularization, direct/solver.rs:838).  solve_many overwrites
 * the n x nrhs column-major panel b with A^-1 b (solve_many,
 * direct/mod.rs:29-41).  `_dev` variants take device pointers. */
typedef struct sdpx_gpu_dense sdpx_gpu_dense;
int32_t sdpx_gpu_dense_create(uint64_t n, sdpx_gpu_dense **out);
int32_t sdpx_gpu_dense_destroy(sdpx_gpu_dense *h);
/* Dynamic pivot rule of the dense factor (dense_block.rs:329-365, settings
 * dynamic_regularization_eps / _delta): a pivot <= eps becomes delta unless
 * it is below -64 * eps_f64 * 64 * |unfactored diagonal| (then info > 0). */
int32_t sdpx_gpu_dense_set_dynamic_pivot(sdpx_gpu_dense *h, int32_t enable, double eps,
                                         double delta);
int32_t sdpx_gpu_dense_refactor(sdpx_gpu_dense *h, const double *a, uint64_t lda, int32_t *info);
int32_t sdpx_gpu_dense_refactor_dev(sdpx_gpu_dense *h, const double *a, uint64_t lda,
                                    int32_t *info);
int32_t sdpx_gpu_dense_solve_many(sdpx_gpu_dense *h, double *b, uint64_t nrhs, uint64_t ldb);
int32_t sdpx_gpu_dense_solve_many_dev(sdpx_gpu_dense *h, double *b, uint64_t nrhs, uint64_t ldb);
/* Device address of the lower factor (n x n, ld n) after a successful
 * refactor; valid until the next refactor or destroy. */
int32_t sdpx_gpu_dense_factor_dev_ptr(sdpx_gpu_dense *h, const double **l);
/* Copy the lower factor to host memory (n x n, ldo); the strict upper part
 * of the copy is unspecified. */
int32_t sdpx_gpu_dense_get_factor(sdpx_gpu_dense *h, double *out, uint64_t ldo);

/* Float64 Schur assembly of one PSD block (replaces the dense-column path of
 * PsdBlock::compute_schur_selected, kkt/condensed/psd.rs:1299-1456, with
 * Ginv = Rinv^T Rinv from kkt.rs:1119-1127):
 *   S_ab = <A_a, Ginv A_b Ginv> = svec(Rinv A_a Rinv^T) . svec(Rinv A_b Rinv^T)
 * for m constraint columns given as svec vectors (s = n(n+1)/2 rows,
 * upper-triangular column order, off-diagonals scaled by sqrt 2), Rinv n x n
 * column-major.  Writes the upper triangle of S (m x m, lds). */
int32_t sdpx_gpu_schur_psd_f64(uint64_t n, uint64_t m, const double *rinv, const double *a_svec,
                               double *s, uint64_t lds);
int32_t sdpx_gpu_schur_psd_f64_dev(uint64_t n, uint64_t m, const double *rinv,
                                   const double *a_svec, double *s, uint64_t lds);

/* Nesterov-Todd scaling of a batch of PSD blocks in binary64, device
 * pointers except dims and status (host): replaces
 * PSDTriangleCone::update_scaling (crates/solver/src/solver/cones/
 * psdtrianglecone.rs:311-430) for every PSD block of the composite cone at
 * once (CompositeCone::update_scaling, cones/compositecone.rs:1603).  Block i
 * of order dims[i] <= 256 reads its slack s and dual z as svec'd upper
 * triangles (off-diagonals times sqrt(2); svec_to_mat, algebra/dense/
 * matrix_math.rs:156-179), concatenated block after block, and writes
 *   R    = L1 V diag(lambda)^-1/2       (n x n, column-major)
 *   Rinv = diag(lambda)^1/2 (L1^-T V)^T (Rinv R = I)
 *   lambda, the singular values of L2^T L1, descending
 *   G    = R R^T                         (both triangles)
 * with L1 = chol(S), L2 = chol(Z) and V the right singular vectors; the
 * matrices are concatenated (sum n^2), lambda too (sum n).  status[i]: 0 ok,
 * 1 S not positive definite, 2 Z not positive definite (both: the
 * reference's update_scaling returns false), 3 the singular value sweeps did
 * not converge.  Agrees with the reference's binary64 path to rounding
 * (LAPACK's summation and SVD algorithm are not specified bit for bit). */
int32_t sdpx_gpu_psd_nt_scaling_f64_dev(uint64_t nblocks, const uint64_t *dims, const double *s, const double *z,
                                        double *r, double *rinv, double *lambda, double *g, int32_t *status);

/* Condensed Schur assembly of a reduced KKT with several cones (replaces
 * CondensedKKT::assemble, kkt/condensed/mod.rs:1962-2147, for Float64
 * PSD and orthant cones; device pointers): the upper triangle of S (m x m,
 * lds) becomes
 *   P (CSC: colptr m + 1, rowval, nzval; entries i <= j; optional)
 *   + each PSD block's S_ab = <A_a, W A_b W> (W = Rinv^T Rinv, the block's
 *     NT scaling: kkt.rs:1119-1127, psd.rs:1299-1456) at its global
 *     columns, blocks published in cone order (2071-2086)
 *   + the orthant rows' Gram B^T diag(w) B (rows x m, ldb; the eliminated
 *     rows of 2090-2101).
 * Summation order differs from the reference's BLAS; parity is by
 * tolerance (1e-12 relative).  struct_size must be sizeof the struct. */
typedef struct {
    uint64_t struct_size;       /* sizeof(sdpx_gpu_psd_block_f64) */
    uint64_t n, ncols;          /* block order, constraint columns touching it */
    const double *rinv;         /* n x n */
    const double *a_svec;       /* s x ncols, s = n(n+1)/2 */
    const uint64_t *columns;    /* ncols global column indices < m */
} sdpx_gpu_psd_block_f64;
int32_t sdpx_gpu_condensed_assemble_f64_dev(uint64_t m, const int64_t *p_colptr, const int64_t *p_rowval,
                                            const double *p_nzval, uint64_t nblocks,
                                            const sdpx_gpu_psd_block_f64 *blocks, uint64_t rows, const double *b,
                                            uint64_t ldb, const double *w, double *s, uint64_t lds);

/* The condensed KKT's right-hand side and recovery of the eliminated rows,
 * over the same blocks, orthant rows b and weights w as
 * sdpx_gpu_condensed_assemble_f64_dev (device pointers).  H^-1 is each
 * cone's inverse scaling of the eliminated rows: svec(W smat(v) W), W =
 * Rinv^T Rinv, for a PSD block (PsdScaling::apply with inverse set,
 * kkt/condensed/psd.rs:1968-2000), w o v for the orthant rows.  bz and z
 * hold the eliminated rows block after block (block i: its n_i (n_i + 1) / 2
 * svec rows), then the `rows` orthant rows.
 *   rhs:     xr = bx + A^T H^-1 bz   (m; replaces CondensedKKT::prepare_rhs,
 *            kkt/condensed/mod.rs:2416-2507)
 *   recover: z = H^-1 (A x - bz)     (replaces CondensedKKT::recover_rhs,
 *            kkt/condensed/mod.rs:2511-2603, without retained rows)
 * so that x solving S x = xr with S the assembled matrix, and z, solve the
 * unreduced system P x + A^T z = bx, A x - H z = bz.  Float64; parity by
 * tolerance. */
int32_t sdpx_gpu_condensed_rhs_f64_dev(uint64_t m, uint64_t nblocks, const sdpx_gpu_psd_block_f64 *blocks,
                                       uint64_t rows, const double *b, uint64_t ldb, const double *w,
                                       const double *bx, const double *bz, double *xr);
int32_t sdpx_gpu_condensed_recover_f64_dev(uint64_t m, uint64_t nblocks, const sdpx_gpu_psd_block_f64 *blocks,
                                           uint64_t rows, const double *b, uint64_t ldb, const double *w,
                                           const double *x, const double *bz, double *z);

/* ---- Adjoint (differentiation) ---------------------------------------
 *
 * No counterpart in the reference (it has no differentiation); built on its
 * KKT solve_many (kkt/mod.rs:59, direct/mod.rs:29-41) and the symmetric
 * quasi-definite KKT (paper/sdpx.tex:389-396), whose transpose is itself.
 * At an IPM solution of the PSD-block problem with the Schur factor in h
 * (M_ij = <A_i, W A_j W>, W = Rinv^T Rinv), for nrhs cotangents ybar of the
 * free variables y* (m x nrhs, column-major):
 *   bgrad = M^-1 ybar                      (m x nrhs) = d(ybar . y*)/db
 *   cgrad = svec(W smat(A bgrad) W)        (s x nrhs) = d(ybar . y*)/d svec(C)
 * with the same GPU factor (one transposed-KKT solve per cotangent).
 * rinv n x n, a_svec s x m (s = n(n+1)/2) as sdpx_gpu_schur_psd_f64. */
int32_t sdpx_gpu_adjoint_psd_f64(sdpx_gpu_dense *h, uint64_t n, const double *rinv,
                                 const double *a_svec, uint64_t nrhs, const double *ybar,
                                 double *bgrad, double *cgrad);
int32_t sdpx_gpu_adjoint_psd_f64_dev(sdpx_gpu_dense *h, uint64_t n, const double *rinv,
                                     const double *a_svec, uint64_t nrhs, const double *ybar,
                                     double *bgrad, double *cgrad);

/* Full adjoint at the solution's central-path point (X, y, S: A(X) = b,
 * A*(y) + S = C, X S = mu I; the final iterate of an IPM that targets mu),
 * with the same factor of M: for nrhs functionals
 *   L_k = ybar_k . y + <Xbar_k, X> + obj_k <C, X>
 * (each of ybar m x nrhs, xbar s x nrhs, obj nrhs optional: null is zero)
 * it returns d L_k / db (bgrad m x nrhs), d L_k / d svec(C) (cgrad s x
 * nrhs) and, when agrad is not null, d L_k / d svec(A_i) for every i
 * (agrad s x m x nrhs, column i of block k).  With W = Rinv^T Rinv the NT
 * scaling (W S W = X), the point moves as dX = -W dS W, dS = dC - dA*(y) -
 * A*(dy), M dy = db + A(W (dC - dA*(y)) W) - dA(X); so with Z = W (Xbar +
 * obj C) W, v = M^-1 (ybar + A(Z)) and U = W A*(v) W:
 *   bgrad = v,  cgrad = U - Z + obj svec(X),  agrad_i = y_i (Z - U) - v_i svec(X).
 * The transposed KKT solve is the same factor (M is symmetric; KKT
 * quasi-definite, paper/sdpx.tex:389-396).  y (m) and x_svec, c_svec (s)
 * are the point's y, svec(X) and svec(C).  No reference counterpart. */
int32_t sdpx_gpu_adjoint_psd_full_f64(sdpx_gpu_dense *h, uint64_t n, const double *rinv, const double *a_svec,
                                      const double *y, const double *x_svec, const double *c_svec, uint64_t nrhs,
                                      const double *ybar, const double *xbar, const double *obj, double *bgrad,
                                      double *cgrad, double *agrad);
int32_t sdpx_gpu_adjoint_psd_full_f64_dev(sdpx_gpu_dense *h, uint64_t n, const double *rinv, const double *a_svec,
                                          const double *y, const double *x_svec, const double *c_svec,
                                          uint64_t nrhs, const double *ybar, const double *xbar,
                                          const double *obj, double *bgrad, double *cgrad, double *agrad);

#ifdef __cplusplus
}
#endif
#endif
