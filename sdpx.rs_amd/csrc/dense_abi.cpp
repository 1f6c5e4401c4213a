// C ABI of the Float64 dense factor/solve and PSD Schur assembly.
#include <algorithm>
#include <cstring>
#include <functional>

#include "common.h"
#include "fp64_dense.h"

struct sdpx_gpu_dense {
    int64_t n = 0;
    double *L = nullptr;   // device n x n, lower factor after refactor
    int *info = nullptr;   // device
    double *ws = nullptr;  // split-K workspace
    size_t ws_bytes = 0;
    double *diag = nullptr;  // device copy of the unfactored diagonal
    bool factored = false;
    bool rule = false;
    double eps = 0, delta = 0;
};

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);
}  // namespace sdpx

using sdpx::abi_boundary;
using sdpx::abi_invalid;

namespace {

int32_t refactor_impl(sdpx_gpu_dense *h, const double *a, uint64_t lda, bool dev, int32_t *info) {
    if (!h) return abi_invalid("null handle");
    if (lda < (uint64_t)h->n) return abi_invalid("lda below n");
    sdpx::Device &d = sdpx::device();
    std::lock_guard<std::recursive_mutex> g(d.mu);
    hipStream_t s = d.stream;
    const int64_t n = h->n;
    if (n == 0) {
        if (info) *info = 0;
        h->factored = true;
        return SDPX_GPU_OK;
    }
    HIP_CHECK(hipMemcpy2DAsync(h->L, n * 8, a, lda * 8, n * 8, n,
                               dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    sdpx::mirror_upper(h->L, n, n, s);
    HIP_CHECK(hipMemsetAsync(h->info, 0, sizeof(int), s));
    sdpx::PivotRule rule{h->eps, h->delta, h->diag};
    if (h->rule)
        HIP_CHECK(hipMemcpy2DAsync(h->diag, 8, h->L, (n + 1) * 8, 8, n, hipMemcpyDeviceToDevice, s));
    sdpx::potrf_f64(h->L, n, n, h->info, h->ws, h->ws_bytes, s, h->rule ? &rule : nullptr);
    int hinfo = 0;
    HIP_CHECK(hipMemcpyAsync(&hinfo, h->info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (info) *info = hinfo;
    h->factored = hinfo == 0;
    return SDPX_GPU_OK;
}

int32_t solve_impl(sdpx_gpu_dense *h, double *b, uint64_t nrhs, uint64_t ldb, bool dev) {
    if (!h) return abi_invalid("null handle");
    if (!h->factored) return abi_invalid("no successful factorization");
    if (ldb < (uint64_t)h->n) return abi_invalid("ldb below n");
    sdpx::Device &d = sdpx::device();
    std::lock_guard<std::recursive_mutex> g(d.mu);
    hipStream_t s = d.stream;
    const int64_t n = h->n;
    if (n == 0 || nrhs == 0) return SDPX_GPU_OK;
    double *db = b;
    int64_t ld = (int64_t)ldb;
    if (!dev) {
        db = (double *)d.ws.get(40, sizeof(double) * n * nrhs);
        ld = n;
        HIP_CHECK(hipMemcpy2DAsync(db, n * 8, b, ldb * 8, n * 8, nrhs, hipMemcpyHostToDevice, s));
    }
    sdpx::potrs_f64(h->L, n, n, db, (int64_t)nrhs, ld, h->ws, h->ws_bytes, s);
    if (!dev)
        HIP_CHECK(hipMemcpy2DAsync(b, ldb * 8, db, n * 8, n * 8, nrhs, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return SDPX_GPU_OK;
}

int32_t schur_impl(uint64_t n, uint64_t m, const double *rinv, const double *a_svec, double *S,
                   uint64_t lds, bool dev) {
    if (lds < m) return abi_invalid("lds below m");
    if (n == 0 || m == 0) return SDPX_GPU_OK;
    sdpx::Device &d = sdpx::device();
    std::lock_guard<std::recursive_mutex> g(d.mu);
    hipStream_t s = d.stream;
    const uint64_t sv = n * (n + 1) / 2;
    const double *dr = rinv, *da = a_svec;
    double *dS = S;
    int64_t ld = (int64_t)lds;
    if (!dev) {
        double *r = (double *)d.ws.get(41, 8 * n * n);
        double *a = (double *)d.ws.get(42, 8 * sv * m);
        dS = (double *)d.ws.get(43, 8 * m * m);
        ld = (int64_t)m;
        HIP_CHECK(hipMemcpyAsync(r, rinv, 8 * n * n, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(a, a_svec, 8 * sv * m, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpy2DAsync(dS, m * 8, S, lds * 8, m * 8, m, hipMemcpyHostToDevice, s));
        dr = r;
        da = a;
    }
    sdpx::schur_psd_f64((int64_t)n, (int64_t)m, dr, da, dS, ld, s);
    if (!dev)
        HIP_CHECK(hipMemcpy2DAsync(S, lds * 8, dS, m * 8, m * 8, m, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return SDPX_GPU_OK;
}

}  // namespace

extern "C" {

int32_t sdpx_gpu_dense_create(uint64_t n, sdpx_gpu_dense **out) {
    return abi_boundary([&]() -> int32_t {
        if (!out) return abi_invalid("null output handle");
        if (n > (1u << 20)) return abi_invalid("order too large");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        auto *h = new sdpx_gpu_dense();
        h->n = (int64_t)n;
        try {
            HIP_CHECK(hipMalloc(&h->L, 8 * std::max<uint64_t>(n * n, 1)));
            HIP_CHECK(hipMalloc(&h->info, sizeof(int)));
            h->ws_bytes = 8 * std::max<uint64_t>(n * n, 1) * 8;
            HIP_CHECK(hipMalloc(&h->ws, h->ws_bytes));
            HIP_CHECK(hipMalloc(&h->diag, 8 * std::max<uint64_t>(n, 1)));
        } catch (...) {
            if (h->L) (void)hipFree(h->L);
            if (h->info) (void)hipFree(h->info);
            delete h;
            throw;
        }
        *out = h;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_dense_destroy(sdpx_gpu_dense *h) {
    return abi_boundary([&]() -> int32_t {
        if (!h) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        (void)hipFree(h->L);
        (void)hipFree(h->info);
        (void)hipFree(h->ws);
        (void)hipFree(h->diag);
        delete h;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_dense_set_dynamic_pivot(sdpx_gpu_dense *h, int32_t enable, double eps,
                                         double delta) {
    return abi_boundary([&]() -> int32_t {
        if (!h) return abi_invalid("null handle");
        if (enable && !(delta > 0.0)) return abi_invalid("delta must be positive");
        h->rule = enable != 0;
        h->eps = eps;
        h->delta = delta;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_dense_refactor(sdpx_gpu_dense *h, const double *a, uint64_t lda, int32_t *info) {
    return abi_boundary([&] { return refactor_impl(h, a, lda, false, info); });
}
int32_t sdpx_gpu_dense_refactor_dev(sdpx_gpu_dense *h, const double *a, uint64_t lda,
                                    int32_t *info) {
    return abi_boundary([&] { return refactor_impl(h, a, lda, true, info); });
}
int32_t sdpx_gpu_dense_solve_many(sdpx_gpu_dense *h, double *b, uint64_t nrhs, uint64_t ldb) {
    return abi_boundary([&] { return solve_impl(h, b, nrhs, ldb, false); });
}
int32_t sdpx_gpu_dense_solve_many_dev(sdpx_gpu_dense *h, double *b, uint64_t nrhs, uint64_t ldb) {
    return abi_boundary([&] { return solve_impl(h, b, nrhs, ldb, true); });
}
int32_t sdpx_gpu_dense_factor_dev_ptr(sdpx_gpu_dense *h, const double **l) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !l) return abi_invalid("null argument");
        *l = h->L;
        return SDPX_GPU_OK;
    });
}
int32_t sdpx_gpu_dense_get_factor(sdpx_gpu_dense *h, double *out, uint64_t ldo) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !out) return abi_invalid("null argument");
        if (!h->factored) return abi_invalid("no successful factorization");
        if (ldo < (uint64_t)h->n) return abi_invalid("ldo below n");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        if (h->n)
            HIP_CHECK(hipMemcpy2DAsync(out, ldo * 8, h->L, h->n * 8, h->n * 8, h->n,
                                       hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_schur_psd_f64(uint64_t n, uint64_t m, const double *rinv, const double *a_svec,
                               double *s, uint64_t lds) {
    return abi_boundary([&] { return schur_impl(n, m, rinv, a_svec, s, lds, false); });
}
int32_t sdpx_gpu_schur_psd_f64_dev(uint64_t n, uint64_t m, const double *rinv,
                                   const double *a_svec, double *s, uint64_t lds) {
    return abi_boundary([&] { return schur_impl(n, m, rinv, a_svec, s, lds, true); });
}

}  // extern "C"

#include "fp64_sampled.h"

extern "C" {

int32_t sdpx_gpu_sampled_arrow_f64_dev(uint64_t nleaves, const sdpx_gpu_sampled_leaf_f64 *leaves,
                                       uint64_t t, const double *c, double *s, double *border_factor,
                                       double dyn_eps, double dyn_delta, int32_t *info,
                                       double *times_ms) {
    return abi_boundary([&]() -> int32_t {
        if (nleaves && !leaves) return abi_invalid("null leaves");
        if (t && (!c || !s || !border_factor)) return abi_invalid("null border");
        for (uint64_t i = 0; i < nleaves; i++)
            if (leaves[i].struct_size != sizeof(sdpx_gpu_sampled_leaf_f64))
                return abi_invalid("leaf struct_size is not sizeof(sdpx_gpu_sampled_leaf_f64)");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        std::vector<sdpx::Fp64Leaf> in(nleaves);
        size_t hb = 0;
        for (uint64_t i = 0; i < nleaves; i++) hb += leaves[i].samples * leaves[i].samples;
        double *hbuf = (double *)d.ws.get(54, 8 * std::max<size_t>(hb, 1));
        size_t off = 0;
        for (uint64_t i = 0; i < nleaves; i++) {
            const auto &l = leaves[i];
            in[i] = sdpx::Fp64Leaf{(int64_t)l.side, (int64_t)l.samples, l.rinv, l.basis,
                                   l.weights, l.coupling, hbuf + off};
            off += l.samples * l.samples;
        }
        int *dinfo = (int *)d.ws.get(55, sizeof(int));
        HIP_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int), d.stream));
        sdpx::Fp64SampledTimes tm;
        sdpx::PivotRule rule{dyn_eps, dyn_delta, nullptr};
        sdpx::sampled_arrow_f64(in, (int64_t)t, c, s, border_factor, dinfo, &tm, d.stream,
                                dyn_delta > 0 ? &rule : nullptr);
        int hinfo = 0;
        HIP_CHECK(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        if (info) *info = hinfo;
        if (times_ms) {
            times_ms[0] = tm.build_ms;
            times_ms[1] = tm.leaves_ms;
            times_ms[2] = tm.border_ms;
        }
        return SDPX_GPU_OK;
    });
}

}  // extern "C"

extern "C" {

int32_t sdpx_gpu_adjoint_psd_f64(sdpx_gpu_dense *h, uint64_t n, const double *rinv,
                                 const double *a_svec, uint64_t nrhs, const double *ybar,
                                 double *bgrad, double *cgrad) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !h->factored) return abi_invalid("no successful factorization");
        const uint64_t m = (uint64_t)h->n, s = n * (n + 1) / 2;
        if (m == 0 || nrhs == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        hipStream_t st = d.stream;
        double *dr = (double *)d.ws.get(63, 8 * n * n);
        double *da = (double *)d.ws.get(64, 8 * s * m);
        double *dy = (double *)d.ws.get(65, 8 * m * nrhs);
        double *db = (double *)d.ws.get(66, 8 * m * nrhs);
        double *dc = (double *)d.ws.get(67, 8 * s * nrhs);
        HIP_CHECK(hipMemcpyAsync(dr, rinv, 8 * n * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(da, a_svec, 8 * s * m, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(dy, ybar, 8 * m * nrhs, hipMemcpyHostToDevice, st));
        sdpx::adjoint_psd_f64(h->L, (int64_t)m, (int64_t)n, dr, da, (int64_t)nrhs, dy, db, dc, st);
        HIP_CHECK(hipMemcpyAsync(bgrad, db, 8 * m * nrhs, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(cgrad, dc, 8 * s * nrhs, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_adjoint_psd_f64_dev(sdpx_gpu_dense *h, uint64_t n, const double *rinv,
                                     const double *a_svec, uint64_t nrhs, const double *ybar,
                                     double *bgrad, double *cgrad) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !h->factored) return abi_invalid("no successful factorization");
        if (h->n == 0 || nrhs == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::adjoint_psd_f64(h->L, h->n, (int64_t)n, rinv, a_svec, (int64_t)nrhs, ybar, bgrad,
                              cgrad, d.stream);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}


int32_t sdpx_gpu_adjoint_psd_full_f64_dev(sdpx_gpu_dense *h, uint64_t n, const double *rinv, const double *a_svec,
                                          const double *y, const double *x_svec, const double *c_svec,
                                          uint64_t nrhs, const double *ybar, const double *xbar,
                                          const double *obj, double *bgrad, double *cgrad, double *agrad) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !h->factored) return abi_invalid("no successful factorization");
        if (!rinv || !a_svec || !y || !x_svec || !bgrad || !cgrad) return abi_invalid("null operand");
        if (obj && !c_svec) return abi_invalid("an objective weight needs C");
        if (h->n == 0 || nrhs == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::adjoint_psd_full_f64(h->L, h->n, (int64_t)n, rinv, a_svec, y, x_svec, c_svec, (int64_t)nrhs, ybar,
                                   xbar, obj, bgrad, cgrad, agrad, d.stream);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_adjoint_psd_full_f64(sdpx_gpu_dense *h, uint64_t n, const double *rinv, const double *a_svec,
                                      const double *y, const double *x_svec, const double *c_svec, uint64_t nrhs,
                                      const double *ybar, const double *xbar, const double *obj, double *bgrad,
                                      double *cgrad, double *agrad) {
    return abi_boundary([&]() -> int32_t {
        if (!h || !h->factored) return abi_invalid("no successful factorization");
        if (!rinv || !a_svec || !y || !x_svec || !bgrad || !cgrad) return abi_invalid("null operand");
        if (obj && !c_svec) return abi_invalid("an objective weight needs C");
        const uint64_t m = (uint64_t)h->n, s = n * (n + 1) / 2;
        if (m == 0 || nrhs == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        hipStream_t st = d.stream;
        auto put = [&](int slot, const double *src, size_t count) -> double * {
            if (!src) return nullptr;
            double *p = (double *)d.ws.get(slot, 8 * std::max<size_t>(count, 1));
            HIP_CHECK(hipMemcpyAsync(p, src, 8 * count, hipMemcpyHostToDevice, st));
            return p;
        };
        // staging slots of the full adjoint (100-110: no other entry point uses them)
        double *dr = put(100, rinv, n * n), *da = put(101, a_svec, s * m), *dy = put(102, y, m);
        double *dx = put(103, x_svec, s), *dc = put(104, c_svec, s), *dyb = put(105, ybar, m * nrhs);
        double *dxb = put(106, xbar, s * nrhs), *dob = put(107, obj, nrhs);
        double *db = (double *)d.ws.get(108, 8 * m * nrhs), *dcg = (double *)d.ws.get(109, 8 * s * nrhs);
        double *dag = agrad ? (double *)d.ws.get(110, 8 * s * m * nrhs) : nullptr;
        sdpx::adjoint_psd_full_f64(h->L, (int64_t)m, (int64_t)n, dr, da, dy, dx, dc, (int64_t)nrhs, dyb, dxb, dob,
                                   db, dcg, dag, st);
        HIP_CHECK(hipMemcpyAsync(bgrad, db, 8 * m * nrhs, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(cgrad, dcg, 8 * s * nrhs, hipMemcpyDeviceToHost, st));
        if (agrad) HIP_CHECK(hipMemcpyAsync(agrad, dag, 8 * s * m * nrhs, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        return SDPX_GPU_OK;
    });
}


int32_t sdpx_gpu_condensed_assemble_f64_dev(uint64_t m, const int64_t *p_colptr, const int64_t *p_rowval,
                                            const double *p_nzval, uint64_t nblocks,
                                            const sdpx_gpu_psd_block_f64 *blocks, uint64_t rows, const double *b,
                                            uint64_t ldb, const double *w, double *s, uint64_t lds) {
    return abi_boundary([&]() -> int32_t {
        if (lds < m) return abi_invalid("lds below m");
        if (m && !s) return abi_invalid("null output");
        if ((p_colptr == nullptr) != (p_nzval == nullptr) || (p_colptr && !p_rowval))
            return abi_invalid("P needs colptr, rowval and nzval");
        if (rows && (!b || !w || ldb < rows)) return abi_invalid("orthant rows need b (ldb >= rows) and w");
        if (nblocks && !blocks) return abi_invalid("null blocks");
        std::vector<sdpx::PsdBlockF64> bl(nblocks);
        for (uint64_t i = 0; i < nblocks; i++) {
            const auto &x = blocks[i];
            if (x.struct_size != sizeof(sdpx_gpu_psd_block_f64)) return abi_invalid("psd block struct_size");
            if (x.ncols && (!x.rinv || !x.a_svec || !x.columns)) return abi_invalid("null block array");
            bl[i] = sdpx::PsdBlockF64{(int64_t)x.n, (int64_t)x.ncols, x.rinv, x.a_svec, x.columns};
        }
        if (m == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::condensed_assemble_f64((int64_t)m, p_colptr, p_rowval, p_nzval, bl, (int64_t)rows, b, (int64_t)ldb, w,
                                     s, (int64_t)lds, d.stream);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

// The blocks of a condensed call, checked (shared by the rhs and recovery).
static int32_t condensed_blocks(uint64_t nblocks, const sdpx_gpu_psd_block_f64 *blocks,
                                std::vector<sdpx::PsdBlockF64> &bl, uint64_t &eliminated) {
    if (nblocks && !blocks) return abi_invalid("null blocks");
    bl.resize(nblocks);
    eliminated = 0;
    for (uint64_t i = 0; i < nblocks; i++) {
        const auto &x = blocks[i];
        if (x.struct_size != sizeof(sdpx_gpu_psd_block_f64)) return abi_invalid("psd block struct_size");
        if (!x.rinv || (x.ncols && (!x.a_svec || !x.columns))) return abi_invalid("null block array");
        bl[i] = sdpx::PsdBlockF64{(int64_t)x.n, (int64_t)x.ncols, x.rinv, x.a_svec, x.columns};
        eliminated += x.n * (x.n + 1) / 2;
    }
    return SDPX_GPU_OK;
}

int32_t sdpx_gpu_condensed_rhs_f64_dev(uint64_t m, uint64_t nblocks, const sdpx_gpu_psd_block_f64 *blocks,
                                       uint64_t rows, const double *b, uint64_t ldb, const double *w,
                                       const double *bx, const double *bz, double *xr) {
    return abi_boundary([&]() -> int32_t {
        if (rows && (!b || !w || ldb < rows)) return abi_invalid("orthant rows need b (ldb >= rows) and w");
        std::vector<sdpx::PsdBlockF64> bl;
        uint64_t el = 0;
        const int32_t r = condensed_blocks(nblocks, blocks, bl, el);
        if (r != SDPX_GPU_OK) return r;
        if (m && (!bx || !xr)) return abi_invalid("null bx or xr");
        if (el + rows && !bz) return abi_invalid("null bz");
        if (m == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::condensed_rhs_f64((int64_t)m, bl, (int64_t)rows, b, (int64_t)ldb, w, bx, bz, xr, d.stream);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_condensed_recover_f64_dev(uint64_t m, uint64_t nblocks, const sdpx_gpu_psd_block_f64 *blocks,
                                           uint64_t rows, const double *b, uint64_t ldb, const double *w,
                                           const double *x, const double *bz, double *z) {
    return abi_boundary([&]() -> int32_t {
        if (rows && (!b || !w || ldb < rows)) return abi_invalid("orthant rows need b (ldb >= rows) and w");
        std::vector<sdpx::PsdBlockF64> bl;
        uint64_t el = 0;
        const int32_t r = condensed_blocks(nblocks, blocks, bl, el);
        if (r != SDPX_GPU_OK) return r;
        if (el + rows == 0) return SDPX_GPU_OK;
        if (!bz || !z || (m && !x)) return abi_invalid("null x, bz or z");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::condensed_recover_f64((int64_t)m, bl, (int64_t)rows, b, (int64_t)ldb, w, x, bz, z, d.stream);
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

}  // extern "C"

#include "nt_scaling.h"

extern "C" {

int32_t sdpx_gpu_psd_nt_scaling_f64_dev(uint64_t nblocks, const uint64_t *dims, const double *s, const double *z,
                                        double *r, double *rinv, double *lambda, double *g, int32_t *status) {
    return abi_boundary([&]() -> int32_t {
        if (nblocks && (!dims || !status)) return abi_invalid("null dims or status");
        std::vector<sdpx::NtBlock> small, large;
        std::vector<size_t> small_idx, large_idx;
        int64_t so = 0, mo = 0, vo = 0, smax = 0, lmax = 0;
        for (uint64_t i = 0; i < nblocks; i++) {
            const int64_t n = (int64_t)dims[i];
            if (n > sdpx::NT_MAX_ORDER) return abi_invalid("PSD block order above 256");
            const sdpx::NtBlock b{n, so, mo, vo};
            if (n == 0) {
                status[i] = 0;
            } else if (n <= 64) {
                small.push_back(b);
                small_idx.push_back(i);
                smax = std::max(smax, n);
            } else {
                large.push_back(b);
                large_idx.push_back(i);
                lmax = std::max(lmax, n);
            }
            so += n * (n + 1) / 2;
            mo += n * n;
            vo += n;
        }
        if (so && (!s || !z || !r || !rinv || !lambda || !g)) return abi_invalid("null block arrays");
        if (small.empty() && large.empty()) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> guard(d.mu);
        if (!small.empty()) {
            sdpx::NtBlock *db = (sdpx::NtBlock *)d.ws.get(170, sizeof(sdpx::NtBlock) * small.size());
            int *dst = (int *)d.ws.get(171, sizeof(int) * small.size());
            HIP_CHECK(hipMemcpyAsync(db, small.data(), sizeof(sdpx::NtBlock) * small.size(), hipMemcpyHostToDevice,
                                     d.stream));
            sdpx::nt_scaling_f64(db, (int)small.size(), smax, s, z, r, rinv, lambda, g, dst, 60, d.stream);
            std::vector<int> hs(small.size());
            HIP_CHECK(hipMemcpyAsync(hs.data(), dst, sizeof(int) * hs.size(), hipMemcpyDeviceToHost, d.stream));
            HIP_CHECK(hipStreamSynchronize(d.stream));
            for (size_t i = 0; i < small.size(); i++) status[small_idx[i]] = hs[i];
        }
        if (!large.empty()) {
            double *work = (double *)d.ws.get(172, sdpx::nt_scaling_large_workspace_bytes(lmax));
            for (size_t i = 0; i < large.size(); i++) {
                const sdpx::NtBlock &b = large[i];
                status[large_idx[i]] = sdpx::nt_scaling_large_f64(b.n, s + b.svec_off, z + b.svec_off, r + b.mat_off,
                                                                  rinv + b.mat_off, lambda + b.vec_off,
                                                                  g + b.mat_off, work, 60, d.stream);
            }
            HIP_CHECK(hipStreamSynchronize(d.stream));
        }
        return SDPX_GPU_OK;
    });
}

}  // extern "C"
