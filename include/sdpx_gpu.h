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
 * sdpx_gpu_stream) and return once the results are complete.
 */
#ifndef SDPX_GPU_H
#define SDPX_GPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SDPX_GPU_ABI_VERSION 1u
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
int32_t sdpx_gpu_last_stage_times(double *ms, uint32_t count);

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

#ifdef __cplusplus
}
#endif
#endif
