// The C ABI boundary: status codes, thread-local last error, host staging.
// Mirrors crates/ffi/src/lib.rs:30-43 (boundary(): errors become codes and a
// thread-local message; nothing unwinds across the ABI).
#include <cstring>
#include <functional>
#include <string>

#include "common.h"
#include "mp_kernels.h"
#include "mpx.h"
#include "rns_gemm.h"

namespace {
thread_local std::string t_error;

template <class F>
int32_t boundary(F &&f) {
    try {
        int32_t code = f();
        if (code != SDPX_GPU_OK && code != SDPX_GPU_DECLINED) {
            if (t_error.empty()) t_error = "error";
        }
        return code;
    } catch (const sdpx::Error &e) {
        t_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        t_error = std::string("internal error: ") + e.what();
        return SDPX_GPU_DEVICE_ERROR;
    } catch (...) {
        t_error = "internal error";
        return SDPX_GPU_DEVICE_ERROR;
    }
}

int32_t invalid(const char *what) {
    t_error = what;
    return SDPX_GPU_INVALID;
}
}  // namespace

namespace sdpx {
int32_t check_exact_blocks(uint64_t m, uint64_t nblocks, const sdpx_gpu_exact_block *blocks, uint64_t &K);
int32_t blocks_upper_device(int limbs, uint64_t m, uint64_t nblocks, const sdpx_gpu_exact_block *blocks, void *out,
                            bool dev);
int32_t abi_boundary(const std::function<int32_t()> &f) { return boundary(f); }
int32_t abi_invalid(const char *what) { return invalid(what); }
void set_last_error(const char *what) { t_error = what; }
}  // namespace sdpx

namespace {

bool trans_ok(uint8_t t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
char trans_norm(uint8_t t) { return (t == 'N' || t == 'n') ? 'N' : 'T'; }

// Span (in values) a column-major rows x cols view with leading dim ld covers.
uint64_t span(uint64_t rows, uint64_t cols, uint64_t ld) {
    return rows == 0 || cols == 0 ? 0 : (cols - 1) * ld + rows;
}

int32_t check_gemm(uint32_t limbs, uint8_t ta, uint8_t tb, uint64_t m, uint64_t n, uint64_t k,
                   const void *a, uint64_t lda, const void *b, uint64_t ldb, uint32_t upper,
                   void *out, uint64_t out_len) {
    if (limbs < 1 || limbs > (uint32_t)sdpx::MAX_LIMBS) return invalid("limbs must be 1..16");
    if (!trans_ok(ta) || !trans_ok(tb)) return invalid("transpose must be N or T");
    uint64_t ar = trans_norm(ta) == 'N' ? m : k, br = trans_norm(tb) == 'N' ? k : n;
    if ((ar > 0 && lda < ar) || (br > 0 && ldb < br)) return invalid("leading dimension too small");
    if ((m * k > 0 && !a) || (k * n > 0 && !b) || (m * n > 0 && !out)) return invalid("null pointer");
    bool packed = upper && m == n && out_len == n * (n + 1) / 2;
    if (!packed && out_len < m * n) return invalid("output length below m*n");
    if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) return invalid("dimension too large");
    return SDPX_GPU_OK;
}
}  // namespace

extern "C" {

uint32_t sdpx_gpu_abi_version(void) { return SDPX_GPU_ABI_VERSION; }

int32_t sdpx_gpu_init(int32_t dev) {
    return boundary([&] {
        sdpx::device_select(dev);
        return SDPX_GPU_OK;
    });
}

void *sdpx_gpu_stream(void) {
    void *s = nullptr;
    boundary([&] {
        s = (void *)sdpx::device().stream;
        return SDPX_GPU_OK;
    });
    return s;
}

uint64_t sdpx_gpu_last_error(char *out, uint64_t capacity) {
    uint64_t need = t_error.size() + 1;
    if (out && capacity > 0) {
        uint64_t nb = need <= capacity ? need - 1 : capacity - 1;
        std::memcpy(out, t_error.data(), nb);
        out[nb] = 0;
    }
    return need;
}

int32_t sdpx_gpu_last_plan(uint32_t *moduli, uint32_t *planes, double *bits) {
    const sdpx::RnsStats &s = sdpx::rns_last_stats();
    if (moduli) *moduli = (uint32_t)s.moduli;
    if (planes) *planes = (uint32_t)s.planes;
    if (bits) *bits = s.bits;
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_set_tuning(const char *name, int64_t value) {
    if (!name || !sdpx::mp_set_tuning(name, value)) return invalid("unknown tuning switch");
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_set_profiling(int32_t on) {
    sdpx::rns_set_profiling(on != 0);
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_last_stage_times(double *ms, uint32_t count) {
    const sdpx::RnsStats &s = sdpx::rns_last_stats();
    for (uint32_t i = 0; i < count && i < 4; i++) ms[i] = s.stage_ms[i];
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_rns_totals(int32_t reset, double *stage_ms4, double *int8_ops, double *mp_macs,
                            int64_t *calls) {
    sdpx::RnsTotals &t = sdpx::rns_totals();
    if (stage_ms4)
        for (int i = 0; i < 4; i++) stage_ms4[i] = t.stage_ms[i];
    if (int8_ops) *int8_ops = t.int8_ops;
    if (mp_macs) *mp_macs = t.mp_macs;
    if (calls) *calls = t.calls;
    if (reset) t = sdpx::RnsTotals{};
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_rns_gemm_dev(uint32_t limbs, uint8_t ta, uint8_t tb, uint64_t m, uint64_t n,
                              uint64_t k, const void *a, uint64_t lda, const void *b, uint64_t ldb,
                              uint32_t upper, void *out, uint64_t out_len) {
    return boundary([&]() -> int32_t {
        int32_t c = check_gemm(limbs, ta, tb, m, n, k, a, lda, b, ldb, upper, out, out_len);
        if (c) return c;
        sdpx::Device &dev = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dev.mu);
        sdpx::RnsGemmArgs args{(int)limbs, trans_norm(ta), trans_norm(tb), (int64_t)m, (int64_t)n,
                               (int64_t)k, a, (int64_t)lda, b, (int64_t)ldb, upper != 0, out_len,
                               out};
        const int32_t r = sdpx::rns_gemm_device(args);
        HIP_CHECK(hipStreamSynchronize(dev.stream));  // the result is complete on return
        return r;
    });
}

int32_t sdpx_gpu_rns_gemm_dev_noalloc(uint32_t limbs, uint8_t ta, uint8_t tb, uint64_t m, uint64_t n,
                                      uint64_t k, const void *a, uint64_t lda, const void *b,
                                      uint64_t ldb, uint32_t upper, void *out, uint64_t out_len) {
    return boundary([&]() -> int32_t {
        int32_t c = check_gemm(limbs, ta, tb, m, n, k, a, lda, b, ldb, upper, out, out_len);
        if (c) return c;
        sdpx::Device &dev = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dev.mu);
        sdpx::RnsGemmArgs args{(int)limbs, trans_norm(ta), trans_norm(tb), (int64_t)m, (int64_t)n,
                               (int64_t)k, a, (int64_t)lda, b, (int64_t)ldb, upper != 0, out_len,
                               out};
        args.scratch = sdpx::rns_noalloc_scratch();
        args.scratch_bytes = sdpx::NOALLOC_SCRATCH_BYTES;
        const int32_t r = sdpx::rns_gemm_device(args);
        HIP_CHECK(hipStreamSynchronize(dev.stream));  // the result is complete on return
        return r;
    });
}

int32_t sdpx_gpu_rns_gemm(uint32_t limbs, uint8_t ta, uint8_t tb, uint64_t m, uint64_t n,
                          uint64_t k, const void *a, uint64_t lda, const void *b, uint64_t ldb,
                          uint32_t upper, void *out, uint64_t out_len) {
    return boundary([&]() -> int32_t {
        int32_t c = check_gemm(limbs, ta, tb, m, n, k, a, lda, b, ldb, upper, out, out_len);
        if (c) return c;
        sdpx::Device &dev = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(dev.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        uint64_t sa = span(trans_norm(ta) == 'N' ? m : k, trans_norm(ta) == 'N' ? k : m, lda);
        uint64_t sb = span(trans_norm(tb) == 'N' ? k : n, trans_norm(tb) == 'N' ? n : k, ldb);
        bool packed = upper && m == n && out_len == n * (n + 1) / 2;
        uint64_t so = packed ? out_len : m * n;
        // staging slots above the product's own
        void *da = dev.ws.get(20, sa * vb);
        void *db = (a == b && sa == sb) ? da : dev.ws.get(21, sb * vb);
        void *dout = dev.ws.get(22, so * vb);
        hipStream_t s = dev.stream;
        if (sa) HIP_CHECK(hipMemcpyAsync(da, a, sa * vb, hipMemcpyHostToDevice, s));
        if (sb && db != da) HIP_CHECK(hipMemcpyAsync(db, b, sb * vb, hipMemcpyHostToDevice, s));
        // untouched outputs (strict lower triangle of unpacked upper) keep their values
        if (upper && !packed && so)
            HIP_CHECK(hipMemcpyAsync(dout, out, so * vb, hipMemcpyHostToDevice, s));
        sdpx::RnsGemmArgs args{(int)limbs, trans_norm(ta), trans_norm(tb), (int64_t)m, (int64_t)n,
                               (int64_t)k, da, (int64_t)lda, db, (int64_t)ldb, upper != 0,
                               out_len, dout};
        int32_t r = sdpx::rns_gemm_device(args);
        if (r == SDPX_GPU_OK && so) {
            HIP_CHECK(hipMemcpyAsync(out, dout, so * vb, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return r;
    });
}

int32_t sdpx_gpu_rns_gemm_blocks_upper(uint32_t limbs, uint64_t m, uint64_t nblocks,
                                       const sdpx_gpu_exact_block *blocks, void *out) {
    return boundary([&]() -> int32_t {
        if (limbs < 1 || limbs > (uint32_t)sdpx::MAX_LIMBS) return invalid("limbs must be 1..16");
        const size_t vb = sdpx::value_bytes((int)limbs);
        uint64_t K = 0;
        if (int32_t r = sdpx::check_exact_blocks(m, nblocks, blocks, K)) return r;
        if (m == 0) return SDPX_GPU_OK;
        if (!out) return invalid("null output");
        if (sdpx::mp_kernels_support((int)limbs))
            return sdpx::blocks_upper_device((int)limbs, m, nblocks, blocks, out, false);
        // (other widths) scatter every block into dense K x m operands Z (= RN(diag(d) y)) and Y;
        // columns a block does not touch are exact zeros, which leave every
        // exact sum unchanged (rns_blas.rs:2671-2922 sums the blocks exactly
        // and rounds once)
        std::vector<uint8_t> Z(K * m * vb), Y(K * m * vb);
        for (size_t e = 0; e < K * m; e++) {
            sdpx::mpx::set_zero(&Z[e * vb], (int)limbs);
            sdpx::mpx::set_zero(&Y[e * vb], (int)limbs);
        }
        uint64_t r0 = 0;
        for (uint64_t b = 0; b < nblocks; b++) {
            const auto &bl = blocks[b];
            const uint8_t *y = (const uint8_t *)bl.y, *d = (const uint8_t *)bl.d;
            for (uint64_t c = 0; c < bl.ncols; c++)
                for (uint64_t r = 0; r < bl.rows; r++) {
                    const uint8_t *src = y + (r + c * bl.rows) * vb;
                    const size_t dst = (r0 + r) + bl.columns[c] * K;
                    sdpx::mpx::copy_value(&Y[dst * vb], src, (int)limbs);
                    sdpx::mpx::mul(sdpx::mpx::View{(int)limbs, d + r * vb},
                                   sdpx::mpx::View{(int)limbs, src}, &Z[dst * vb]);
                }
            r0 += bl.rows;
        }
        // c[i + j m] (i <= j) = sum_r Z[r, i] Y[r, j]
        return sdpx_gpu_rns_gemm(limbs, 'T', 'N', m, m, K, Z.data(), K, Y.data(), K, 1, out, m * m);
    });
}

}  // extern "C"
