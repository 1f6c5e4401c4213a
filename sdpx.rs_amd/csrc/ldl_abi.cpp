// C ABI of the multiple-precision dense LDL^T.
#include <omp.h>

#include <cstring>
#include <functional>

#include "common.h"
#include "ldl.h"
#include "ldl_device.h"
#include "mp_kernels.h"

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);
void set_last_error(const char *what);
}  // namespace sdpx

namespace {
thread_local sdpx::LdlStats t_ldl_stats;
}

extern "C" {

int32_t sdpx_gpu_mp_ldl(uint32_t limbs, uint64_t n, void *a, uint64_t lda, const int8_t *signs,
                        uint64_t nsigns, const void *reg_eps, const void *reg_delta, void *dinv,
                        uint32_t flags, int32_t *ok) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (limbs < 1 || limbs > (uint32_t)sdpx::MAX_LIMBS) return sdpx::abi_invalid("limbs must be 1..16");
        if (n > 0 && (!a || !dinv)) return sdpx::abi_invalid("null pointer");
        if (lda < n) return sdpx::abi_invalid("lda below n");
        if (nsigns != 0 && nsigns != 1 && nsigns != n) return sdpx::abi_invalid("signs: 0, 1 or n entries");
        if ((reg_eps == nullptr) != (reg_delta == nullptr))
            return sdpx::abi_invalid("reg_eps and reg_delta go together");
        sdpx::LdlOptions opt;
        opt.threads = omp_get_max_threads();
        opt.gpu_products = (flags & SDPX_GPU_LDL_HOST_PRODUCTS) == 0;
        opt.reg = reg_eps != nullptr;
        opt.reg_eps = (const uint8_t *)reg_eps;
        opt.reg_delta = (const uint8_t *)reg_delta;
        bool recursive = (flags & SDPX_GPU_LDL_DOTS) == 0 && n >= 256;
        t_ldl_stats = sdpx::LdlStats{};
        if (opt.gpu_products && sdpx::mp_kernels_support((int)limbs) && n > 0) {
            // device-resident factor: stage in, factor in HBM, stage out
            sdpx::Device &d = sdpx::device();
            std::lock_guard<std::recursive_mutex> g(d.mu);
            const size_t vb = sdpx::value_bytes((int)limbs);
            uint8_t *dl = (uint8_t *)d.ws.get(72, n * n * vb);
            uint8_t *dd = (uint8_t *)d.ws.get(73, n * vb);
            HIP_CHECK(hipMemcpy2DAsync(dl, n * vb, a, lda * vb, n * vb, n, hipMemcpyHostToDevice, d.stream));
            HIP_CHECK(hipStreamSynchronize(d.stream));
            int r = sdpx::ldl_factor_device((int)limbs, dl, n, n, signs, nsigns, reg_eps, reg_delta,
                                            dd, t_ldl_stats, recursive);
            if (r >= 0) {
                if (r == 1) {
                    HIP_CHECK(hipMemcpy2DAsync(a, lda * vb, dl, n * vb, n * vb, n, hipMemcpyDeviceToHost,
                                               d.stream));
                    HIP_CHECK(hipMemcpyAsync(dinv, dd, n * vb, hipMemcpyDeviceToHost, d.stream));
                    HIP_CHECK(hipStreamSynchronize(d.stream));
                }
                if (ok) *ok = r;
                return SDPX_GPU_OK;
            }
            t_ldl_stats.declined++;  // not handled on the GPU: host path below
        }
        bool good = sdpx::ldl_factor((int)limbs, (uint8_t *)a, lda, n, signs, nsigns, opt,
                                     (uint8_t *)dinv, t_ldl_stats, recursive);
        if (ok) *ok = good ? 1 : 0;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_ldl_dev(uint32_t limbs, uint64_t n, void *a, uint64_t lda, const int8_t *signs,
                            uint64_t nsigns, const void *reg_eps, const void *reg_delta, void *dinv,
                            uint32_t flags, int32_t *ok) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("device factor supports 128, 256 and 512 bits");
        if (lda < n) return sdpx::abi_invalid("lda below n");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        t_ldl_stats = sdpx::LdlStats{};
        bool recursive = (flags & SDPX_GPU_LDL_DOTS) == 0 && n >= 256;
        int r = sdpx::ldl_factor_device((int)limbs, (uint8_t *)a, lda, n, signs, nsigns, reg_eps,
                                        reg_delta, (uint8_t *)dinv, t_ldl_stats, recursive);
        if (r < 0) return SDPX_GPU_DECLINED;
        if (ok) *ok = r;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_ldl_solve(uint32_t limbs, uint64_t n, const void *l, uint64_t ldl,
                              const void *dinv, uint64_t nrhs, void *b, uint64_t ldb) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("GPU solves support 128, 256 and 512 bits");
        if (ldl < n || ldb < n) return sdpx::abi_invalid("leading dimension below n");
        if (n == 0 || nrhs == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        const size_t vb = sdpx::value_bytes((int)limbs);
        uint8_t *dl = (uint8_t *)d.ws.get(72, n * n * vb);
        uint8_t *dd = (uint8_t *)d.ws.get(73, n * vb);
        uint8_t *db = (uint8_t *)d.ws.get(74, n * nrhs * vb);
        HIP_CHECK(hipMemcpy2DAsync(dl, n * vb, l, ldl * vb, n * vb, n, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipMemcpyAsync(dd, dinv, n * vb, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipMemcpy2DAsync(db, n * vb, b, ldb * vb, n * vb, nrhs, hipMemcpyHostToDevice, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        int r = sdpx::ldl_solve_device((int)limbs, dl, n, n, dd, db, n, nrhs);
        if (r < 0) return SDPX_GPU_DECLINED;
        HIP_CHECK(hipMemcpy2DAsync(b, ldb * vb, db, n * vb, n * vb, nrhs, hipMemcpyDeviceToHost, d.stream));
        HIP_CHECK(hipStreamSynchronize(d.stream));
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_ldl_solve_dev(uint32_t limbs, uint64_t n, const void *l, uint64_t ldl,
                                  const void *dinv, uint64_t nrhs, void *b, uint64_t ldb) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("GPU solves support 128, 256 and 512 bits");
        if (ldl < n || ldb < n) return sdpx::abi_invalid("leading dimension below n");
        if (n == 0 || nrhs == 0) return SDPX_GPU_OK;
        if (!l || !dinv || !b) return sdpx::abi_invalid("null array");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        int r = sdpx::ldl_solve_device((int)limbs, (const uint8_t *)l, ldl, n, (const uint8_t *)dinv,
                                       (uint8_t *)b, ldb, nrhs);
        return r < 0 ? SDPX_GPU_DECLINED : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_ldl_stats(int64_t *gpu_products, int64_t *host_products, int64_t *regularized,
                              double *gpu_mp_macs, int64_t *declined, int64_t *retried) {
    if (gpu_products) *gpu_products = t_ldl_stats.gpu_products;
    if (host_products) *host_products = t_ldl_stats.host_products;
    if (regularized) *regularized = t_ldl_stats.regularized;
    if (gpu_mp_macs) *gpu_mp_macs = t_ldl_stats.gpu_mp_macs;
    if (declined) *declined = t_ldl_stats.declined;
    if (retried) *retried = t_ldl_stats.retried;
    return SDPX_GPU_OK;
}

}  // extern "C"

#include "sampled.h"

namespace {
thread_local sdpx::SampledStats t_sampled_stats;

// Leaves of order (samples) 256 and more factor recursively and solve Y by
// recursive panels in the reference (factor_loaded, arrow.rs:819;
// Leaf::refactor, arrow.rs:1596-1632); this pipeline factors every leaf by
// exact dots, so it declines them (nothing written; the caller falls back).
int32_t leaf_order_limit(uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves) {
    for (uint64_t i = 0; i < nleaves; i++)
        if (leaves[i].struct_size != sizeof(sdpx_gpu_sampled_leaf))
            return sdpx::abi_invalid("leaf struct_size is not sizeof(sdpx_gpu_sampled_leaf)");
    for (uint64_t i = 0; i < nleaves; i++)
        if (leaves[i].samples >= 256) {
            sdpx::set_last_error("sampled leaf of order 256 or more (recursive leaf factor) declined");
            return SDPX_GPU_DECLINED;
        }
    return 0;
}

// Stage a host-resident sampled problem through HBM and run the device
// pipeline; -1 when the GPU path declines (then the host path runs).
int sampled_host_via_device(int limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                            uint64_t t, uint8_t *border, const int8_t *bsigns, uint64_t nbsigns,
                            uint8_t *border_dinv, sdpx_gpu_sampled_factor *factors,
                            sdpx::SampledStats &st) {
    sdpx::Device &d = sdpx::device();
    std::lock_guard<std::recursive_mutex> g(d.mu);
    hipStream_t s = d.stream;
    const size_t vb = sdpx::value_bytes(limbs);
    size_t in_vals = 0, lsum = 0, ksum = 0;
    for (uint64_t i = 0; i < nleaves; i++) {
        const auto &l = leaves[i];
        in_vals += l.side * l.side + l.side * l.samples + l.samples + l.samples * t;
        lsum += l.samples * l.samples;
        ksum += l.samples;
    }
    uint8_t *din = (uint8_t *)d.ws.get(74, std::max<size_t>(in_vals, 1) * vb);
    uint8_t *dl = (uint8_t *)d.ws.get(75, std::max<size_t>(lsum, 1) * vb);
    uint8_t *dd = (uint8_t *)d.ws.get(76, std::max<size_t>(ksum, 1) * vb);
    uint8_t *dy = (uint8_t *)d.ws.get(77, std::max<size_t>(ksum * t, 1) * vb);
    uint8_t *db = (uint8_t *)d.ws.get(78, std::max<size_t>(t * t, 1) * vb);
    uint8_t *dbd = (uint8_t *)d.ws.get(79, std::max<size_t>(t, 1) * vb);
    std::vector<sdpx::SampledLeafDev> dev(nleaves);
    size_t off = 0;
    auto put = [&](const void *src, size_t count) {
        uint8_t *p = din + off * vb;
        if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * vb, hipMemcpyHostToDevice, s));
        off += count;
        return p;
    };
    for (uint64_t i = 0; i < nleaves; i++) {
        const auto &l = leaves[i];
        dev[i].h = l.side;
        dev[i].k = l.samples;
        dev[i].rinv = put(l.rinv, l.side * l.side);
        dev[i].q = put(l.basis, l.side * l.samples);
        dev[i].w = put(l.weights, l.samples);
        dev[i].b = put(l.coupling, l.samples * t);
    }
    if (t) HIP_CHECK(hipMemcpyAsync(db, border, t * t * vb, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    int r = sdpx::sampled_arrow_device(limbs, dev, t, dl, dd, dy, db, bsigns, nbsigns, dbd, st);
    if (r < 0) return -1;
    if (t) {
        HIP_CHECK(hipMemcpyAsync(border, db, t * t * vb, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(border_dinv, dbd, t * vb, hipMemcpyDeviceToHost, s));
    }
    if (factors) {
        size_t lo = 0, ko = 0;
        for (uint64_t i = 0; i < nleaves; i++) {
            const size_t K = leaves[i].samples;
            auto &f = factors[i];
            if (f.l) HIP_CHECK(hipMemcpyAsync(f.l, dl + lo * vb, K * K * vb, hipMemcpyDeviceToHost, s));
            if (f.dinv) HIP_CHECK(hipMemcpyAsync(f.dinv, dd + ko * vb, K * vb, hipMemcpyDeviceToHost, s));
            if (f.y) HIP_CHECK(hipMemcpyAsync(f.y, dy + ko * t * vb, K * t * vb, hipMemcpyDeviceToHost, s));
            lo += K * K;
            ko += K;
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    return r;
}
}  // namespace

extern "C" {

int32_t sdpx_gpu_sampled_arrow_dev(uint32_t limbs, uint64_t nleaves,
                                   const sdpx_gpu_sampled_leaf *leaves, uint64_t t, void *l_all,
                                   void *dinv_all, void *y_all, void *border,
                                   const int8_t *border_signs, uint64_t nborder_signs,
                                   void *border_dinv, int32_t *ok) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("device pipeline supports 128, 256 and 512 bits");
        if (nleaves && !leaves) return sdpx::abi_invalid("null leaves");
        if (t && (!border || !border_dinv)) return sdpx::abi_invalid("null border");
        if (nborder_signs != 0 && nborder_signs != 1 && nborder_signs != t)
            return sdpx::abi_invalid("border signs: 0, 1 or t entries");
        if (nborder_signs && !border_signs) return sdpx::abi_invalid("null border signs");
        if (int32_t r = leaf_order_limit(nleaves, leaves)) return r;
        for (uint64_t i = 0; i < nleaves; i++)
            if (!leaves[i].rinv || !leaves[i].basis || !leaves[i].weights ||
                (t && !leaves[i].coupling))
                return sdpx::abi_invalid("null leaf array");
        if (nleaves && (!l_all || !dinv_all || (t && !y_all)))
            return sdpx::abi_invalid("null leaf factor output");
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        std::vector<sdpx::SampledLeafDev> dev(nleaves);
        for (uint64_t i = 0; i < nleaves; i++)
            dev[i] = sdpx::SampledLeafDev{leaves[i].side, leaves[i].samples,
                                          (const uint8_t *)leaves[i].rinv,
                                          (const uint8_t *)leaves[i].basis,
                                          (const uint8_t *)leaves[i].weights,
                                          (const uint8_t *)leaves[i].coupling};
        t_sampled_stats = sdpx::SampledStats{};
        int r = sdpx::sampled_arrow_device((int)limbs, dev, t, (uint8_t *)l_all, (uint8_t *)dinv_all,
                                           (uint8_t *)y_all, (uint8_t *)border, border_signs,
                                           nborder_signs, (uint8_t *)border_dinv, t_sampled_stats);
        if (r < 0) return SDPX_GPU_DECLINED;
        if (ok) *ok = r;
        return SDPX_GPU_OK;
    });
}

}  // extern "C"

extern "C" {

int32_t sdpx_gpu_sampled_arrow(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                               uint64_t t, void *border, const int8_t *border_signs,
                               uint64_t nborder_signs, void *border_dinv,
                               sdpx_gpu_sampled_factor *factors, uint32_t flags, int32_t *ok) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (limbs < 1 || limbs > (uint32_t)sdpx::MAX_LIMBS) return sdpx::abi_invalid("limbs must be 1..16");
        if (nleaves && !leaves) return sdpx::abi_invalid("null leaves");
        if (t && (!border || !border_dinv)) return sdpx::abi_invalid("null border");
        if (nborder_signs != 0 && nborder_signs != 1 && nborder_signs != t)
            return sdpx::abi_invalid("border signs: 0, 1 or t entries");
        if (int32_t r = leaf_order_limit(nleaves, leaves)) return r;
        if (factors)
            for (uint64_t i = 0; i < nleaves; i++)
                if (factors[i].struct_size != sizeof(sdpx_gpu_sampled_factor))
                    return sdpx::abi_invalid("factor struct_size is not sizeof(sdpx_gpu_sampled_factor)");
        std::vector<sdpx::SampledLeafIn> in(nleaves);
        for (uint64_t i = 0; i < nleaves; i++) {
            const auto &l = leaves[i];
            if (!l.rinv || !l.basis || !l.weights || (t && !l.coupling))
                return sdpx::abi_invalid("null leaf array");
            in[i] = sdpx::SampledLeafIn{l.side, l.samples, (const uint8_t *)l.rinv,
                                        (const uint8_t *)l.basis, (const uint8_t *)l.weights,
                                        (const uint8_t *)l.coupling};
        }
        sdpx::LdlOptions opt;
        opt.threads = omp_get_max_threads();
        opt.gpu_products = (flags & SDPX_GPU_LDL_HOST_PRODUCTS) == 0;
        std::vector<sdpx::SampledLeafOut> out;
        t_sampled_stats = sdpx::SampledStats{};
        if (opt.gpu_products && sdpx::mp_kernels_support((int)limbs)) {
            int r = sampled_host_via_device((int)limbs, nleaves, leaves, t, (uint8_t *)border,
                                            border_signs, nborder_signs, (uint8_t *)border_dinv,
                                            factors, t_sampled_stats);
            if (r >= 0) {
                if (ok) *ok = r;
                return SDPX_GPU_OK;
            }
            t_sampled_stats.declined++;  // not handled on the GPU: host panels below
        }
        bool good = sdpx::sampled_arrow_factor((int)limbs, in, t, (uint8_t *)border, border_signs,
                                               nborder_signs, opt, out, (uint8_t *)border_dinv,
                                               t_sampled_stats);
        if (factors) {
            const size_t vb = sdpx::value_bytes((int)limbs);
            for (uint64_t i = 0; i < nleaves && i < out.size(); i++) {
                auto &f = factors[i];
                const size_t K = leaves[i].samples;
                if (f.l && out[i].l.size() == K * K * vb) std::memcpy(f.l, out[i].l.data(), K * K * vb);
                if (f.dinv && out[i].dinv.size() == K * vb) std::memcpy(f.dinv, out[i].dinv.data(), K * vb);
                if (f.y && out[i].y.size() == K * t * vb) std::memcpy(f.y, out[i].y.data(), K * t * vb);
            }
        }
        if (ok) *ok = good ? 1 : 0;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_sampled_arrow_solve_dev(uint32_t limbs, uint64_t nleaves, const uint64_t *samples, uint64_t t,
                                         const void *l_all, const void *dinv_all, const void *y_all,
                                         const void *border, const void *border_dinv, uint64_t nrhs, void *x,
                                         uint64_t ldx) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("arrow solves support 128, 256 and 512 bits");
        if (nleaves && !samples) return sdpx::abi_invalid("null leaf orders");
        uint64_t n = t;
        std::vector<size_t> ks(nleaves);
        for (uint64_t i = 0; i < nleaves; i++) {
            ks[i] = samples[i];
            n += samples[i];
        }
        if (ldx < n) return sdpx::abi_invalid("ldx below the system order");
        if (nrhs && n && !x) return sdpx::abi_invalid("null right-hand sides");
        if (nleaves && (!l_all || !dinv_all || (t && !y_all))) return sdpx::abi_invalid("null leaf factor");
        if (t && (!border || !border_dinv)) return sdpx::abi_invalid("null border factor");
        if (nrhs == 0 || n == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        const int r = sdpx::sampled_arrow_solve_device((int)limbs, ks, t, (const uint8_t *)l_all,
                                                       (const uint8_t *)dinv_all, (const uint8_t *)y_all,
                                                       (const uint8_t *)border, (const uint8_t *)border_dinv, nrhs,
                                                       (uint8_t *)x, ldx);
        return r < 0 ? SDPX_GPU_DECLINED : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_sampled_schur_dev(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                                   void *h_all) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("the sampled Schur build supports 128, 256 and 512 bits");
        if (nleaves && (!leaves || !h_all)) return sdpx::abi_invalid("null pointer");
        std::vector<sdpx::SampledLeafDev> in(nleaves);
        for (uint64_t i = 0; i < nleaves; i++) {
            const auto &l = leaves[i];
            if (l.struct_size != sizeof(sdpx_gpu_sampled_leaf))
                return sdpx::abi_invalid("leaf struct_size is not sizeof(sdpx_gpu_sampled_leaf)");
            if (!l.rinv || !l.basis || !l.weights) return sdpx::abi_invalid("null leaf array");
            in[i] = sdpx::SampledLeafDev{l.side, l.samples, (const uint8_t *)l.rinv, (const uint8_t *)l.basis,
                                         (const uint8_t *)l.weights, (const uint8_t *)l.coupling};
        }
        if (nleaves == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        return sdpx::sampled_schur_device((int)limbs, in, (uint8_t *)h_all) < 0 ? SDPX_GPU_DECLINED : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_sampled_kkt_solve_dev(uint32_t limbs, uint64_t nleaves, const sdpx_gpu_sampled_leaf *leaves,
                                       uint64_t t, const void *h_all, const void *c, const void *l_all,
                                       const void *dinv_all, const void *y_all, const void *border,
                                       const void *border_dinv, const sdpx_gpu_refine_settings *settings,
                                       uint64_t nrhs, const void *b, uint64_t ldb, void *x, uint64_t ldx,
                                       int32_t *ok, void *norms, int64_t *refinements) {
    return sdpx::abi_boundary([&]() -> int32_t {
        if (!sdpx::mp_kernels_support((int)limbs))
            return sdpx::abi_invalid("the refined arrow solve supports 128, 256 and 512 bits");
        if (!settings || settings->struct_size != sizeof(sdpx_gpu_refine_settings))
            return sdpx::abi_invalid("settings struct_size is not sizeof(sdpx_gpu_refine_settings)");
        if (!settings->reltol || !settings->abstol || !settings->stop_ratio)
            return sdpx::abi_invalid("null refinement tolerance");
        if (nleaves && !leaves) return sdpx::abi_invalid("null leaves");
        uint64_t n = t;
        std::vector<size_t> ks(nleaves);
        std::vector<const uint8_t *> coupling(nleaves);
        for (uint64_t i = 0; i < nleaves; i++) {
            if (leaves[i].struct_size != sizeof(sdpx_gpu_sampled_leaf))
                return sdpx::abi_invalid("leaf struct_size is not sizeof(sdpx_gpu_sampled_leaf)");
            if (t && !leaves[i].coupling) return sdpx::abi_invalid("null leaf coupling");
            ks[i] = leaves[i].samples;
            coupling[i] = (const uint8_t *)leaves[i].coupling;
            n += leaves[i].samples;
        }
        if (ldb < n || ldx < n) return sdpx::abi_invalid("ldb or ldx below the system order");
        if (nrhs && n && (!b || !x)) return sdpx::abi_invalid("null right-hand sides or solutions");
        if (nleaves && (!h_all || !l_all || !dinv_all || (t && !y_all))) return sdpx::abi_invalid("null leaf array");
        if (t && (!c || !border || !border_dinv)) return sdpx::abi_invalid("null border array");
        if (refinements) *refinements = 0;
        if (nrhs == 0 || n == 0) return SDPX_GPU_OK;
        sdpx::Device &d = sdpx::device();
        std::lock_guard<std::recursive_mutex> g(d.mu);
        sdpx::RefineSettings rs;
        rs.reltol = (const uint8_t *)settings->reltol;
        rs.abstol = (const uint8_t *)settings->abstol;
        rs.stop_ratio = (const uint8_t *)settings->stop_ratio;
        rs.max_iter = settings->max_iter;
        const int r = sdpx::sampled_kkt_solve_device(
            (int)limbs, ks, t, (const uint8_t *)h_all, coupling, (const uint8_t *)c, (const uint8_t *)l_all,
            (const uint8_t *)dinv_all, (const uint8_t *)y_all, (const uint8_t *)border, (const uint8_t *)border_dinv,
            rs, nrhs, (const uint8_t *)b, ldb, (uint8_t *)x, ldx, ok, (uint8_t *)norms, refinements);
        return r < 0 ? SDPX_GPU_DECLINED : SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_panel_phase_times(double *ms4, int64_t *launches4, int32_t reset) {
    return sdpx::abi_boundary([&]() -> int32_t {
        sdpx::panel_phase_times(ms4, launches4, reset != 0);
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_sampled_stats(double *seconds3, int64_t *gpu_products, int64_t *host_products,
                               int64_t *declined, int64_t *retried, int64_t *failed_leaf) {
    if (seconds3) {
        seconds3[0] = t_sampled_stats.build_s;
        seconds3[1] = t_sampled_stats.leaves_s;
        seconds3[2] = t_sampled_stats.border_s;
    }
    if (gpu_products) *gpu_products = t_sampled_stats.gpu_products;
    if (host_products) *host_products = t_sampled_stats.host_products;
    if (declined) *declined = t_sampled_stats.declined;
    if (retried) *retried = t_sampled_stats.retried;
    if (failed_leaf) *failed_leaf = t_sampled_stats.failed_leaf;
    return SDPX_GPU_OK;
}

}  // extern "C"
