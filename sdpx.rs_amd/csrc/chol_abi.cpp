// C ABI of the exact-dot dense Cholesky, triangular solve and solve
// (XpotrfScalar / XpotrsScalar for MPFR, blas/mpfr.rs:726-872, and
// xtrsm_lower, blas/traits.rs:85-113).  Host arrays are staged through HBM;
// every subtracted sum is one exact dot rounded once, as in the reference,
// so the results are bitwise the reference's.
#include <cstring>
#include <functional>

#include "common.h"
#include "mp_kernels.h"

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);
}  // namespace sdpx

namespace {

// Kind word of record idx (records: `limbs` u64 limbs, kind, exponent).
inline int32_t kind_of(const void *base, size_t idx, size_t vb, int limbs) {
    int32_t k;
    std::memcpy(&k, (const uint8_t *)base + idx * vb + 8 * (size_t)limbs, 4);
    return k;
}
inline bool finite_kind(int32_t k) { return k == 2 || k == -2 || k == 3 || k == -3; }

// Every entry of the stored triangle (lower: i >= k at i + k ld) finite.
bool triangle_finite(const void *a, size_t n, size_t ld, bool lower, size_t vb, int limbs) {
    for (size_t c = 0; c < n; c++)
        for (size_t r = lower ? c : 0; r < (lower ? n : c + 1); r++)
            if (!finite_kind(kind_of(a, r + c * ld, vb, limbs))) return false;
    return true;
}
bool block_finite(const void *b, size_t n, size_t ncols, size_t ld, size_t vb, int limbs) {
    for (size_t c = 0; c < ncols; c++)
        for (size_t r = 0; r < n; r++)
            if (!finite_kind(kind_of(b, r + c * ld, vb, limbs))) return false;
    return true;
}
bool valid_uplo(uint8_t u) { return u == 'L' || u == 'l' || u == 'U' || u == 'u'; }
bool lower_uplo(uint8_t u) { return u == 'L' || u == 'l'; }

}  // namespace

extern "C" {

int32_t sdpx_gpu_mp_potrf(uint32_t limbs, uint8_t uplo, uint64_t n, void *a, uint64_t lda, int32_t *info) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!info) return sdpx::abi_invalid("info is null");
        *info = !valid_uplo(uplo) ? -1 : lda < (n > 0 ? n : 1) ? -4 : 0;
        if (*info != 0 || n == 0) return SDPX_GPU_OK;
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("GPU Cholesky supports 128, 256 and 512 bits");
        if (!a) return sdpx::abi_invalid("null array");
        const size_t vb = sdpx::value_bytes((int)limbs);
        const bool lower = lower_uplo(uplo);
        if (!triangle_finite(a, n, lda, lower, vb, (int)limbs)) return SDPX_GPU_DECLINED;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        uint8_t *da = (uint8_t *)d.ws.get(90, n * n * vb);
        int *dst = (int *)d.ws.get(93, 4 * sizeof(int));
        HIP_CHECK(hipMemsetAsync(dst, 0, 4 * sizeof(int), d.stream));
        HIP_CHECK(hipMemcpy2DAsync(da, n * vb, a, lda * vb, n * vb, n, hipMemcpyHostToDevice, d.stream));
        sdpx::mp_potrf((int)limbs, da, (int64_t)n, (int64_t)n, lower ? 1 : 0, dst, d.stream);
        int st[4];
        HIP_CHECK(hipMemcpyAsync(st, dst, sizeof(st), hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        if (st[2]) return SDPX_GPU_DECLINED;  // a dot beyond the exponent window: nothing written
        HIP_CHECK(hipMemcpy2DAsync(a, lda * vb, da, n * vb, n * vb, n, hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        *info = st[0];
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_potrs(uint32_t limbs, uint8_t uplo, uint64_t n, uint64_t nrhs, const void *a,
                          uint64_t lda, void *b, uint64_t ldb, int32_t *info) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!info) return sdpx::abi_invalid("info is null");
        *info = !valid_uplo(uplo) ? -1 : lda < (n > 0 ? n : 1) ? -5 : ldb < (n > 0 ? n : 1) ? -7 : 0;
        if (*info != 0 || n == 0 || nrhs == 0) return SDPX_GPU_OK;
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("GPU solves support 128, 256 and 512 bits");
        if (!a || !b) return sdpx::abi_invalid("null array");
        const size_t vb = sdpx::value_bytes((int)limbs);
        const bool lower = lower_uplo(uplo);
        if (!triangle_finite(a, n, lda, lower, vb, (int)limbs) || !block_finite(b, n, nrhs, ldb, vb, (int)limbs))
            return SDPX_GPU_DECLINED;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        uint8_t *da = (uint8_t *)d.ws.get(90, n * n * vb);
        uint8_t *db = (uint8_t *)d.ws.get(91, n * nrhs * vb);
        HIP_CHECK(hipMemcpy2DAsync(da, n * vb, a, lda * vb, n * vb, n, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipMemcpy2DAsync(db, n * vb, b, ldb * vb, n * vb, nrhs, hipMemcpyHostToDevice, d.stream));
        sdpx::mp_potrs((int)limbs, da, (int64_t)n, (int64_t)n, lower ? 1 : 0, db, (int64_t)n, (int64_t)nrhs,
                       d.stream);
        HIP_CHECK(hipMemcpy2DAsync(b, ldb * vb, db, n * vb, n * vb, nrhs, hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_trsm_lower(uint32_t limbs, uint64_t n, const void *a, void *b, uint64_t nrhs,
                               uint32_t transpose) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (n == 0 || nrhs == 0) return SDPX_GPU_OK;
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("GPU solves support 128, 256 and 512 bits");
        if (!a || !b) return sdpx::abi_invalid("null array");
        const size_t vb = sdpx::value_bytes((int)limbs);
        // the solve reads the lower triangle (and the diagonal) only
        if (!triangle_finite(a, n, n, true, vb, (int)limbs) || !block_finite(b, n, nrhs, n, vb, (int)limbs))
            return SDPX_GPU_DECLINED;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        uint8_t *da = (uint8_t *)d.ws.get(90, n * n * vb);
        uint8_t *db = (uint8_t *)d.ws.get(91, n * nrhs * vb);
        int *dst = (int *)d.ws.get(93, 4 * sizeof(int));
        HIP_CHECK(hipMemsetAsync(dst, 0, 4 * sizeof(int), d.stream));
        HIP_CHECK(hipMemcpyAsync(da, a, n * n * vb, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipMemcpyAsync(db, b, n * nrhs * vb, hipMemcpyHostToDevice, d.stream));
        sdpx::mp_trsm_lower((int)limbs, da, (int64_t)n, db, (int64_t)nrhs, transpose ? 1 : 0, dst, d.stream);
        int st[4];
        HIP_CHECK(hipMemcpyAsync(st, dst, sizeof(st), hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        if (st[0]) return SDPX_GPU_DECLINED;
        HIP_CHECK(hipMemcpyAsync(b, db, n * nrhs * vb, hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

}  // extern "C"
