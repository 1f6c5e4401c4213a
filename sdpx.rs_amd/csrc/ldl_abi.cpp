// C ABI of the multiple-precision dense LDL^T.
#include <omp.h>

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
