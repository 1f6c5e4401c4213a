// C ABI of the multiple-precision dense LDL^T.
#include <omp.h>

#include <cstring>
#include <functional>

#include "common.h"
#include "ldl.h"

namespace sdpx {
int32_t abi_boundary(const std::function<int32_t()> &f);
int32_t abi_invalid(const char *what);
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
        bool good = sdpx::ldl_factor((int)limbs, (uint8_t *)a, lda, n, signs, nsigns, opt,
                                     (uint8_t *)dinv, t_ldl_stats, recursive);
        if (ok) *ok = good ? 1 : 0;
        return SDPX_GPU_OK;
    });
}

int32_t sdpx_gpu_mp_ldl_stats(int64_t *gpu_products, int64_t *host_products, int64_t *regularized,
                              double *gpu_mp_macs) {
    if (gpu_products) *gpu_products = t_ldl_stats.gpu_products;
    if (host_products) *host_products = t_ldl_stats.host_products;
    if (regularized) *regularized = t_ldl_stats.regularized;
    if (gpu_mp_macs) *gpu_mp_macs = t_ldl_stats.gpu_mp_macs;
    return SDPX_GPU_OK;
}

}  // extern "C"

#include "sampled.h"

namespace {
thread_local sdpx::SampledStats t_sampled_stats;
}

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

int32_t sdpx_gpu_sampled_stats(double *seconds3, int64_t *gpu_products, int64_t *host_products) {
    if (seconds3) {
        seconds3[0] = t_sampled_stats.build_s;
        seconds3[1] = t_sampled_stats.leaves_s;
        seconds3[2] = t_sampled_stats.border_s;
    }
    if (gpu_products) *gpu_products = t_sampled_stats.gpu_products;
    if (host_products) *host_products = t_sampled_stats.host_products;
    return SDPX_GPU_OK;
}

}  // extern "C"
