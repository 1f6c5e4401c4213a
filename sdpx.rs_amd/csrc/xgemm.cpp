// General exact products with scaling, the reference's MPFR xgemm and xsyrk
// (crates/solver/src/algebra/dense/blas/mpfr.rs:245-348, 540-626):
//   xgemm  c = axpby(alpha, RN(op(a) op(b)), beta, c)
//   xsyrk  c = axpby(alpha, RN(op(a) op(a)^T), beta, c) on one triangle
// with axpby as mpfr.rs:61-80 (alpha or beta of 0, 1 or -1 taken without a
// product).  The exact product is the residue engine's, rounded once per
// entry (the value the reference's residue path and its per-entry dot_fma
// path both give); alpha = 0 skips it, as there.
#include <algorithm>
#include <cstring>
#include <functional>

#include "common.h"
#include "mp_kernels.h"
#include "rns_gemm.h"

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);

namespace {

enum { XS_A = 140, XS_B, XS_C, XS_AB, XS_P };

// the reference's scalar tests (F == 0, F == 1, F == -1): 0, 1, 2, else 3
int scalar_code(const void *v, int limbs) {
    const uint8_t *p = (const uint8_t *)v;
    int32_t kind, e;
    std::memcpy(&kind, p + 8 * limbs, 4);
    std::memcpy(&e, p + 8 * limbs + 4, 4);
    if (kind == KIND_ZERO || kind == -KIND_ZERO) return 0;
    if (kind != KIND_REGULAR && kind != -KIND_REGULAR) return 3;
    const uint64_t *l = (const uint64_t *)p;
    for (int i = 0; i + 1 < limbs; i++)
        if (l[i]) return 3;
    if (l[limbs - 1] != (1ull << 63) || e != 1) return 3;
    return kind > 0 ? 1 : 2;
}

struct Scalars {
    int acode, bcode;
    void *d;  // alpha, beta on the device
};

Scalars upload_scalars(Device &dv, int limbs, const void *alpha, const void *beta) {
    const size_t vb = value_bytes(limbs);
    Scalars s{scalar_code(alpha, limbs), scalar_code(beta, limbs), dv.ws.get(XS_AB, 2 * vb)};
    upload_async(s.d, alpha, vb);
    upload_async((uint8_t *)s.d + vb, beta, vb);
    return s;
}

// c (device, ldc) = axpby(alpha, RN(op(a) op(b)), beta, c); -1 on a decline
int xgemm_device(int limbs, char ta, char tb, size_t m, size_t n, size_t k, const Scalars &sc, const void *a,
                 size_t lda, const void *b, size_t ldb, void *c, size_t ldc) {
    Device &dv = device();
    const size_t vb = value_bytes(limbs);
    void *P = nullptr;
    if (sc.acode != 0) {
        P = dv.ws.get(XS_P, m * n * vb);
        RnsGemmArgs g{limbs, ta, tb, (int64_t)m, (int64_t)n, (int64_t)k, a, (int64_t)lda, b, (int64_t)ldb, false,
                      (uint64_t)(m * n), P};
        const int r = rns_gemm_device(g);
        if (r == SDPX_GPU_DECLINED) return -1;
        if (r != SDPX_GPU_OK) throw Error(r, "exact product failed");
    }
    mp_axpby(limbs, P, (int64_t)m, (int64_t)m, (int64_t)n, sc.acode, sc.d, sc.bcode, (const uint8_t *)sc.d + vb, c,
             (int64_t)ldc, 0, false, false, dv.stream);
    HIP_CHECK(hipStreamSynchronize(dv.stream));
    return 1;
}

// one triangle of c = axpby(alpha, RN(op(a) op(a)^T), beta, c)
int xsyrk_device(int limbs, bool upper, char t, size_t n, size_t k, const Scalars &sc, const void *a, size_t lda,
                 void *c, size_t ldc, bool packed) {
    Device &dv = device();
    const size_t vb = value_bytes(limbs);
    void *P = nullptr;
    if (sc.acode != 0) {
        P = dv.ws.get(XS_P, n * n * vb);
        RnsGemmArgs g{limbs, t, t == 'N' ? 'T' : 'N', (int64_t)n, (int64_t)n, (int64_t)k, a, (int64_t)lda, a,
                      (int64_t)lda, true, (uint64_t)(n * n), P};
        const int r = rns_gemm_device(g);
        if (r == SDPX_GPU_DECLINED) return -1;
        if (r != SDPX_GPU_OK) throw Error(r, "exact product failed");
    }
    mp_axpby(limbs, P, (int64_t)n, (int64_t)n, (int64_t)n, sc.acode, sc.d, sc.bcode, (const uint8_t *)sc.d + vb, c,
             (int64_t)ldc, upper ? 1 : 2, packed, true, dv.stream);
    HIP_CHECK(hipStreamSynchronize(dv.stream));
    return 1;
}

char trans_of(uint8_t t) { return (t == 'N' || t == 'n') ? 'N' : 'T'; }
bool trans_ok(uint8_t t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
size_t span(size_t rows, size_t cols, size_t ld) { return rows && cols ? ld * (cols - 1) + rows : 0; }

int32_t check_xgemm(uint32_t limbs, uint8_t ta, uint8_t tb, uint64_t m, uint64_t n, uint64_t k, const void *alpha,
                    const void *a, uint64_t lda, const void *b, uint64_t ldb, const void *beta, const void *c,
                    uint64_t ldc) {
    if (!mp_kernels_support((int)limbs)) return abi_invalid("128, 256 and 512 bits");
    if (!trans_ok(ta) || !trans_ok(tb)) return abi_invalid("transpose flags are N or T");
    if (!alpha || !beta) return abi_invalid("null alpha or beta");
    const char A = trans_of(ta), B = trans_of(tb);
    if (lda < std::max<uint64_t>(1, A == 'N' ? m : k) || ldb < std::max<uint64_t>(1, B == 'N' ? k : n) ||
        ldc < std::max<uint64_t>(1, m))
        return abi_invalid("leading dimension too small");
    if (m && n && (!c || (k && (!a || !b)))) return abi_invalid("null operand");
    return 0;
}

int32_t check_xsyrk(uint32_t limbs, uint8_t uplo, uint8_t t, uint64_t n, uint64_t k, const void *alpha,
                    const void *a, uint64_t lda, const void *beta, const void *c, uint64_t ldc, uint64_t c_len,
                    bool &packed) {
    if (!mp_kernels_support((int)limbs)) return abi_invalid("128, 256 and 512 bits");
    if (!trans_ok(t) || !(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l'))
        return abi_invalid("uplo is U or L, trans N or T");
    if (!alpha || !beta) return abi_invalid("null alpha or beta");
    const bool up = uplo == 'U' || uplo == 'u';
    packed = up && ldc == n && c_len == n * (n + 1) / 2;
    if (lda < std::max<uint64_t>(1, trans_of(t) == 'N' ? n : k)) return abi_invalid("leading dimension too small");
    if (!packed && (ldc < std::max<uint64_t>(1, n) || c_len < span(n, n, ldc)))
        return abi_invalid("c too short for n and ldc");
    if (n && (!c || (k && !a))) return abi_invalid("null operand");
    return 0;
}

}  // namespace
}  // namespace sdpx

extern "C" {

int32_t sdpx_gpu_xgemm(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m, uint64_t n, uint64_t k,
                       const void *alpha, const void *a, uint64_t lda, const void *b, uint64_t ldb, const void *beta,
                       void *c, uint64_t ldc) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (int32_t r = sdpx::check_xgemm(limbs, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)) return r;
        if (m == 0 || n == 0) return SDPX_GPU_OK;
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        const char ta = sdpx::trans_of(transa), tb = sdpx::trans_of(transb);
        const sdpx::Scalars sc = sdpx::upload_scalars(dv, (int)limbs, alpha, beta);
        const size_t sa = sc.acode ? sdpx::span(ta == 'N' ? m : k, ta == 'N' ? k : m, lda) : 0;
        const size_t sb = sc.acode ? sdpx::span(tb == 'N' ? k : n, tb == 'N' ? n : k, ldb) : 0;
        const size_t scn = sdpx::span(m, n, ldc);
        void *da = dv.ws.get(sdpx::XS_A, std::max<size_t>(sa, 1) * vb);
        void *db = dv.ws.get(sdpx::XS_B, std::max<size_t>(sb, 1) * vb);
        void *dc = dv.ws.get(sdpx::XS_C, scn * vb);
        if (sa) sdpx::upload_async(da, a, sa * vb);
        if (sb) sdpx::upload_async(db, b, sb * vb);
        sdpx::upload_async(dc, c, scn * vb);
        if (sdpx::xgemm_device((int)limbs, ta, tb, m, n, k, sc, da, lda, db, ldb, dc, ldc) < 0)
            return SDPX_GPU_DECLINED;
        HIP_CHECK(hipMemcpy(c, dc, scn * vb, hipMemcpyDeviceToHost));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_xgemm_dev(uint32_t limbs, uint8_t transa, uint8_t transb, uint64_t m, uint64_t n, uint64_t k,
                           const void *alpha, const void *a, uint64_t lda, const void *b, uint64_t ldb,
                           const void *beta, void *c, uint64_t ldc) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (int32_t r = sdpx::check_xgemm(limbs, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)) return r;
        if (m == 0 || n == 0) return SDPX_GPU_OK;
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const sdpx::Scalars sc = sdpx::upload_scalars(dv, (int)limbs, alpha, beta);
        return sdpx::xgemm_device((int)limbs, sdpx::trans_of(transa), sdpx::trans_of(transb), m, n, k, sc, a, lda, b,
                                  ldb, c, ldc) < 0
                   ? SDPX_GPU_DECLINED
                   : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_xsyrk(uint32_t limbs, uint8_t uplo, uint8_t trans, uint64_t n, uint64_t k, const void *alpha,
                       const void *a, uint64_t lda, const void *beta, void *c, uint64_t ldc, uint64_t c_len) {
    return sdpx::abi_boundary([&]() -> int32_t {
        bool packed = false;
        if (int32_t r = sdpx::check_xsyrk(limbs, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, c_len, packed))
            return r;
        if (n == 0) return SDPX_GPU_OK;
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        const char t = sdpx::trans_of(trans);
        const sdpx::Scalars sc = sdpx::upload_scalars(dv, (int)limbs, alpha, beta);
        const size_t sa = sc.acode ? sdpx::span(t == 'N' ? n : k, t == 'N' ? k : n, lda) : 0;
        const size_t scn = packed ? c_len : sdpx::span(n, n, ldc);
        void *da = dv.ws.get(sdpx::XS_A, std::max<size_t>(sa, 1) * vb);
        void *dc = dv.ws.get(sdpx::XS_C, scn * vb);
        if (sa) sdpx::upload_async(da, a, sa * vb);
        sdpx::upload_async(dc, c, scn * vb);
        if (sdpx::xsyrk_device((int)limbs, uplo == 'U' || uplo == 'u', t, n, k, sc, da, lda, dc, ldc, packed) < 0)
            return SDPX_GPU_DECLINED;
        HIP_CHECK(hipMemcpy(c, dc, scn * vb, hipMemcpyDeviceToHost));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_xsyrk_dev(uint32_t limbs, uint8_t uplo, uint8_t trans, uint64_t n, uint64_t k, const void *alpha,
                           const void *a, uint64_t lda, const void *beta, void *c, uint64_t ldc, uint64_t c_len) {
    return sdpx::abi_boundary([&]() -> int32_t {
        bool packed = false;
        if (int32_t r = sdpx::check_xsyrk(limbs, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, c_len, packed))
            return r;
        if (n == 0) return SDPX_GPU_OK;
        sdpx::Device &dv = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dv.mu);
        const sdpx::Scalars sc = sdpx::upload_scalars(dv, (int)limbs, alpha, beta);
        return sdpx::xsyrk_device((int)limbs, uplo == 'U' || uplo == 'u', sdpx::trans_of(trans), n, k, sc, a, lda, c,
                                  ldc, packed) < 0
                   ? SDPX_GPU_DECLINED
                   : SDPX_GPU_OK;
    });
}

}  // extern "C"
