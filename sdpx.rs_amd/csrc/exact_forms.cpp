// Exact multilinear forms on the INT8 residue engine, each entry the exact
// value rounded once (the reference's residue-BLAS quadratic forms behind
// XgemmScalar, crates/solver/src/algebra/dense/blas/traits.rs:259-296;
// rns_blas.rs:2997-3195, 3563-3790, 4032-4276):
//   diag_congruence    A^T diag(d) A
//   congruence         op(A) X op(A)^T
//   svec_quadratic     v_k = sum_{i<=j} c_ij x_t q_ik q_jk  (c_ii = 1, c_ij = s)
//   symmetric_bilinear q_a^T X q_b for selected pairs
// Two tools keep every form a sum of two-factor products, so one exact
// product of the residue engine rounds it once:
// * a rounded product plus its exact error: d x = RN(d x) + (d x - RN(d x)),
//   both parts representable, so d x enters as two terms;
// * an exact product left as an expansion: the CRT writes the exact integer
//   in P-bit chunks from the top (RnsGemmArgs::split_capacity), whose sum is
//   the exact value, so X op(A)^T enters the outer product as C terms.
// All pointers are device pointers; results land in caller buffers.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "common.h"
#include "mp_kernels.h"
#include "rns_gemm.h"

namespace sdpx {
namespace {

constexpr int MAX_CHUNKS = 8;
// workspace slots of this file
enum { WS_A = 120, WS_B, WS_T, WS_S, WS_G, WS_X, WS_H, WS_L, WS_I };

struct Ctx {
    int limbs;
    size_t vb;
    Device &dev;
    hipStream_t st;
    uint8_t *buf(int slot, size_t values) { return (uint8_t *)dev.ws.get(slot, std::max<size_t>(values, 1) * vb); }
    uint8_t *at(uint8_t *p, size_t i) const { return p + i * vb; }
};

// out (ldc m, upper entries only when upper) = RN(op(a) op(b)); -1 on a decline
int product(Ctx &c, char ta, char tb, size_t m, size_t n, size_t k, const void *a, size_t lda, const void *b,
            size_t ldb, bool upper, void *out) {
    RnsGemmArgs g{c.limbs, ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, a, (int64_t)lda, b, (int64_t)ldb, upper,
                  (uint64_t)(m * n), out};
    const int r = rns_gemm_device(g);
    if (r == SDPX_GPU_DECLINED) return -1;
    if (r != SDPX_GPU_OK) throw Error(r, "exact product failed");
    return 0;
}

// the exact product op(a) op(b) (m x n) as chunks at out + c m n, ld m
int expansion(Ctx &c, char ta, char tb, size_t m, size_t n, size_t k, const void *a, size_t lda, const void *b,
              size_t ldb, void *out, int &chunks) {
    chunks = 0;
    RnsGemmArgs g{c.limbs, ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, a, (int64_t)lda, b, (int64_t)ldb, false,
                  (uint64_t)(m * n), out};
    g.split_capacity = MAX_CHUNKS;
    g.split_used = &chunks;
    const int r = rns_gemm_device(g);
    if (r == SDPX_GPU_DECLINED) return -1;
    if (r != SDPX_GPU_OK) throw Error(r, "exact expansion failed");
    return 0;
}

// dst (times r x cols, ld times r) = src (r x cols, ld lds) stacked `times` times
void stack_rows(Ctx &c, uint8_t *dst, const void *src, size_t lds, size_t r, size_t cols, int times) {
    for (int t = 0; t < times; t++)
        HIP_CHECK(hipMemcpy2DAsync(c.at(dst, t * r), times * r * c.vb, src, lds * c.vb, r * c.vb, cols,
                                   hipMemcpyDeviceToDevice, c.st));
}

// chunks (each r x cols, ld r, at T + c r cols) -> one (C r) x cols matrix (ld C r)
void chunks_to_rows(Ctx &c, uint8_t *dst, const uint8_t *T, size_t r, size_t cols, int C) {
    for (int t = 0; t < C; t++)
        HIP_CHECK(hipMemcpy2DAsync(c.at(dst, t * r), C * r * c.vb, T + t * r * cols * c.vb, r * c.vb, r * c.vb, cols,
                                   hipMemcpyDeviceToDevice, c.st));
}

// an m x m sum of no terms: zeros (upper entries only when upper)
int empty_sum(Ctx &c, size_t m, bool upper, void *out) {
    product(c, 'N', 'N', m, m, 0, nullptr, m, nullptr, 1, upper, out);
    HIP_CHECK(hipStreamSynchronize(c.st));
    return 1;
}

}  // namespace

int diag_congruence_device(int limbs, size_t m, size_t k, const void *a, const void *d, void *out, bool upper) {
    Ctx c{limbs, value_bytes(limbs), device(), device().stream};
    if (m == 0) return 1;
    if (k == 0) return empty_sum(c, m, upper, out);
    // [A; A]^T [hi; lo] with hi + lo = diag(d) A exactly
    uint8_t *A2 = c.buf(WS_A, 2 * k * m), *B2 = c.buf(WS_B, 2 * k * m);
    stack_rows(c, A2, a, k, k, m, 2);
    mp_two_product(limbs, a, (int64_t)k, d, 0, B2, c.at(B2, k), (int64_t)(2 * k), (int64_t)k, (int64_t)m, c.st);
    if (product(c, 'T', 'N', m, m, 2 * k, A2, 2 * k, B2, 2 * k, upper, out) < 0) return -1;
    HIP_CHECK(hipStreamSynchronize(c.st));
    return 1;
}

int congruence_device(int limbs, char ta, size_t m, size_t k, const void *a, size_t lda, const void *x, size_t ldx,
                      void *out, bool upper) {
    Ctx c{limbs, value_bytes(limbs), device(), device().stream};
    if (m == 0) return 1;
    if (k == 0) return empty_sum(c, m, upper, out);
    const bool tan = ta == 'N';
    // T = X op(A)^T (k x m) as an exact expansion
    uint8_t *T = c.buf(WS_T, (size_t)MAX_CHUNKS * k * m);
    int C = 0;
    if (expansion(c, 'N', tan ? 'T' : 'N', k, m, k, x, ldx, a, lda, T, C) < 0) return -1;
    // op(A) [T_0; T_1; ...] with op(A) repeated along the inner dimension
    uint8_t *Ts = c.buf(WS_S, (size_t)C * k * m), *Ar = c.buf(WS_A, (size_t)C * k * m);
    chunks_to_rows(c, Ts, T, k, m, C);
    for (int t = 0; t < C; t++) {
        if (tan)
            HIP_CHECK(hipMemcpy2DAsync(c.at(Ar, t * k * m), m * c.vb, a, lda * c.vb, m * c.vb, k,
                                       hipMemcpyDeviceToDevice, c.st));
        else
            mp_transpose(limbs, a, (int64_t)lda, (int64_t)k, (int64_t)m, c.at(Ar, t * k * m), (int64_t)m, c.st);
    }
    if (product(c, 'N', 'N', m, m, (size_t)C * k, Ar, m, Ts, (size_t)C * k, upper, out) < 0) return -1;
    HIP_CHECK(hipStreamSynchronize(c.st));
    return 1;
}

int symmetric_bilinear_device(int limbs, size_t h, size_t cols, const void *q, const void *x, size_t npairs,
                              const uint64_t *pairs, void *out) {
    Ctx c{limbs, value_bytes(limbs), device(), device().stream};
    if (h == 0) return -1;  // declined, as the reference's form does
    if (npairs == 0) return 1;
    // T = X Q exactly (chunks), then all q_a^T T_b at once, then the pairs
    uint8_t *Xf = c.buf(WS_X, h * h), *T = c.buf(WS_T, (size_t)MAX_CHUNKS * h * cols);
    mp_unpack_symmetric(limbs, x, (int64_t)h, Xf, c.st);
    int C = 0;
    if (expansion(c, 'N', 'N', h, cols, h, Xf, h, q, h, T, C) < 0) return -1;
    uint8_t *Ts = c.buf(WS_S, (size_t)C * h * cols), *Qr = c.buf(WS_A, (size_t)C * h * cols),
            *G = c.buf(WS_G, cols * cols);
    chunks_to_rows(c, Ts, T, h, cols, C);
    stack_rows(c, Qr, q, h, h, cols, C);
    if (product(c, 'T', 'N', cols, cols, (size_t)C * h, Qr, (size_t)C * h, Ts, (size_t)C * h, false, G) < 0)
        return -1;
    std::vector<int64_t> idx(2 * npairs);
    for (size_t p = 0; p < npairs; p++) {
        idx[p] = (int64_t)pairs[2 * p];
        idx[npairs + p] = (int64_t)pairs[2 * p + 1];
    }
    int64_t *di = (int64_t *)c.dev.ws.get(WS_I, sizeof(int64_t) * idx.size());
    HIP_CHECK(hipMemcpyAsync(di, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice, c.st));
    mp_gather(limbs, G, (int64_t)cols, di, di + npairs, (int64_t)npairs, out, c.st);
    HIP_CHECK(hipStreamSynchronize(c.st));
    return 1;
}

int svec_quadratic_device(int limbs, size_t h, size_t kmax, const void *q, bool abs_q, const void *x,
                          const void *sqrt2, void *out) {
    Ctx c{limbs, value_bytes(limbs), device(), device().stream};
    if (kmax == 0) return 1;
    if (h == 0) {
        mp_zero_values(limbs, out, (int64_t)kmax, c.st);
        HIP_CHECK(hipStreamSynchronize(c.st));
        return 1;
    }
    // Q (or |Q|), the weights M = Hi + Lo exactly, T = [Hi Lo] [Q; Q] as
    // chunks, then v = diag(Q^T T)
    uint8_t *Q = c.buf(WS_X, h * kmax);
    HIP_CHECK(hipMemcpyAsync(Q, q, h * kmax * c.vb, hipMemcpyDeviceToDevice, c.st));
    if (abs_q) mp_abs_values(limbs, Q, (int64_t)(h * kmax), c.st);
    uint8_t *M = c.buf(WS_H, 2 * h * h), *Q2 = c.buf(WS_L, 2 * h * kmax);
    mp_quadratic_weights(limbs, x, (int64_t)h, sqrt2, M, c.at(M, h * h), c.st);
    stack_rows(c, Q2, Q, h, h, kmax, 2);
    uint8_t *T = c.buf(WS_T, (size_t)MAX_CHUNKS * h * kmax);
    int C = 0;
    if (expansion(c, 'N', 'N', h, kmax, 2 * h, M, h, Q2, 2 * h, T, C) < 0) return -1;
    uint8_t *Ts = c.buf(WS_S, (size_t)C * h * kmax), *Qr = c.buf(WS_A, (size_t)C * h * kmax),
            *G = c.buf(WS_G, kmax * kmax);
    chunks_to_rows(c, Ts, T, h, kmax, C);
    stack_rows(c, Qr, Q, h, h, kmax, C);
    if (product(c, 'T', 'N', kmax, kmax, (size_t)C * h, Qr, (size_t)C * h, Ts, (size_t)C * h, false, G) < 0)
        return -1;
    std::vector<int64_t> idx(2 * kmax);
    for (size_t k = 0; k < kmax; k++) idx[k] = idx[kmax + k] = (int64_t)k;
    int64_t *di = (int64_t *)c.dev.ws.get(WS_I, sizeof(int64_t) * idx.size());
    HIP_CHECK(hipMemcpyAsync(di, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice, c.st));
    mp_gather(limbs, G, (int64_t)kmax, di, di + kmax, (int64_t)kmax, out, c.st);
    HIP_CHECK(hipStreamSynchronize(c.st));
    return 1;
}

}  // namespace sdpx

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);
}  // namespace sdpx

extern "C" {

int32_t sdpx_gpu_diag_congruence_upper_exact_dev(uint32_t limbs, uint64_t m, uint64_t k, const void *a,
                                                 const void *d, void *c, uint32_t upper_only) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (m && (!c || (k && (!a || !d)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        return sdpx::diag_congruence_device((int)limbs, m, k, a, d, c, upper_only != 0) < 0 ? SDPX_GPU_DECLINED
                                                                                           : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_congruence_exact_dev(uint32_t limbs, uint8_t transa, uint64_t m, uint64_t k, const void *a,
                                      uint64_t lda, const void *x, uint64_t ldx, void *c, uint32_t upper_only) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        const char ta = (transa == 'N' || transa == 'n') ? 'N' : 'T';
        if (ldx < k || lda < (ta == 'N' ? m : k)) return sdpx::abi_invalid("leading dimension too small");
        if (m && (!c || (k && (!a || !x)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        return sdpx::congruence_device((int)limbs, ta, m, k, a, lda, x, ldx, c, upper_only != 0) < 0
                   ? SDPX_GPU_DECLINED
                   : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_svec_quadratic_exact_dev(uint32_t limbs, uint64_t h, uint64_t kmax, const void *q, uint32_t abs_q,
                                          const void *x, const void *sqrt2, void *out) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (kmax && (!out || (h && (!q || !x || !sqrt2)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        return sdpx::svec_quadratic_device((int)limbs, h, kmax, q, abs_q != 0, x, sqrt2, out) < 0 ? SDPX_GPU_DECLINED
                                                                                                : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_symmetric_bilinear_exact_dev(uint32_t limbs, uint64_t h, uint64_t columns, const void *q,
                                              const void *x, uint64_t npairs, const uint64_t *pairs, void *out) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (h == 0) return SDPX_GPU_DECLINED;  // as the reference's form
        if (npairs && (!pairs || !out || !q || !x)) return sdpx::abi_invalid("null operand");
        for (uint64_t p = 0; p < 2 * npairs; p++)
            if (pairs[p] >= columns) return sdpx::abi_invalid("pair index beyond the columns");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        return sdpx::symmetric_bilinear_device((int)limbs, h, columns, q, x, npairs, pairs, out) < 0
                   ? SDPX_GPU_DECLINED
                   : SDPX_GPU_OK;
    });
}

// Host-record forms: operands staged through the library's workspace, the
// same orchestration, results copied back (outputs the form does not write,
// the strict lower triangle of an upper-only result, keep their values).
namespace {
constexpr int ST_A = 130, ST_B = 131, ST_C = 132, ST_D = 133;

void *stage(sdpx::Device &dv, int slot, const void *src, size_t values, size_t vb) {
    void *d = dv.ws.get(slot, std::max<size_t>(values, 1) * vb);
    if (values && src) sdpx::upload_async(d, src, values * vb);
    return d;
}

int32_t fetch(sdpx::Device &dv, int r, void *dst, const void *src, size_t values, size_t vb) {
    if (r < 0) return SDPX_GPU_DECLINED;
    if (values) HIP_CHECK(hipMemcpyAsync(dst, src, values * vb, hipMemcpyDeviceToHost, dv.stream));
    HIP_CHECK(hipStreamSynchronize(dv.stream));
    return SDPX_GPU_OK;
}
}  // namespace

int32_t sdpx_gpu_diag_congruence_upper_exact(uint32_t limbs, uint64_t m, uint64_t k, const void *a, const void *d,
                                             void *c) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (m && (!c || (k && (!a || !d)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        void *da = stage(dv, ST_A, a, k * m, vb), *dd = stage(dv, ST_B, d, k, vb), *dc = stage(dv, ST_C, c, m * m, vb);
        return fetch(dv, sdpx::diag_congruence_device((int)limbs, m, k, da, dd, dc, true), c, dc, m * m, vb);
    });
}

int32_t sdpx_gpu_congruence_exact(uint32_t limbs, uint8_t transa, uint64_t m, uint64_t k, const void *a,
                                  uint64_t lda, const void *x, uint64_t ldx, void *c, uint32_t upper_only) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        const char ta = (transa == 'N' || transa == 'n') ? 'N' : 'T';
        if (ldx < k || lda < (ta == 'N' ? m : k)) return sdpx::abi_invalid("leading dimension too small");
        if (m && (!c || (k && (!a || !x)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        const size_t sa = (m && k) ? (ta == 'N' ? lda * (k - 1) + m : lda * (m - 1) + k) : 0;
        const size_t sx = k ? ldx * (k - 1) + k : 0;
        void *da = stage(dv, ST_A, a, sa, vb), *dx = stage(dv, ST_B, x, sx, vb);
        void *dc = stage(dv, ST_C, upper_only ? c : nullptr, m * m, vb);
        return fetch(dv, sdpx::congruence_device((int)limbs, ta, m, k, da, lda, dx, ldx, dc, upper_only != 0), c,
                     dc, m * m, vb);
    });
}

int32_t sdpx_gpu_svec_quadratic_exact(uint32_t limbs, uint64_t h, uint64_t kmax, const void *q, uint32_t abs_q,
                                      const void *x, const void *sqrt2, void *out) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (kmax && (!out || (h && (!q || !x || !sqrt2)))) return sdpx::abi_invalid("null operand");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        void *dq = stage(dv, ST_A, q, h * kmax, vb), *dx = stage(dv, ST_B, x, h * (h + 1) / 2, vb);
        void *ds = stage(dv, ST_D, sqrt2, 1, vb), *dout = stage(dv, ST_C, nullptr, kmax, vb);
        return fetch(dv, sdpx::svec_quadratic_device((int)limbs, h, kmax, dq, abs_q != 0, dx, ds, dout), out, dout,
                     kmax, vb);
    });
}

int32_t sdpx_gpu_symmetric_bilinear_exact(uint32_t limbs, uint64_t h, uint64_t columns, const void *q,
                                          const void *x, uint64_t npairs, const uint64_t *pairs, void *out) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs)) return sdpx::abi_invalid("128, 256 and 512 bits");
        if (h == 0) return SDPX_GPU_DECLINED;  // as the reference's form
        if (npairs && (!pairs || !out || !q || !x)) return sdpx::abi_invalid("null operand");
        for (uint64_t p = 0; p < 2 * npairs; p++)
            if (pairs[p] >= columns) return sdpx::abi_invalid("pair index beyond the columns");
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        void *dq = stage(dv, ST_A, q, h * columns, vb), *dx = stage(dv, ST_B, x, h * (h + 1) / 2, vb);
        void *dout = stage(dv, ST_C, nullptr, npairs, vb);
        return fetch(dv, sdpx::symmetric_bilinear_device((int)limbs, h, columns, dq, dx, npairs, pairs, dout), out,
                     dout, npairs, vb);
    });
}

}  // extern "C"
