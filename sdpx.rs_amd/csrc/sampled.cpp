// Per-iteration Schur build + arrow factor of a sampled (bootstrap-style)
// SDP in multiple precision, with every exact product on the GPU.
//
// Follows the reference's fully sampled condensed KKT with the arrow LDL^T:
//   V = Rinv Q, Gram = V^T V          sampled/mod.rs:1836-1868, 1990-2030,
//                                      2065-2143 (exact products, rounded once)
//   S_pq = w_p w_q 1/2 (K_pq K_pq + K_pq K_pq)
//                                      sampled/mod.rs:1973-1988 (dim 1), with the
//                                      reference's left-to-right roundings
//   leaf H = LDL^T                     arrow.rs:1522-1560 -> factor_dots
//                                      (leaves below order 256), 156-253
//   Y = L^-1 B                         forward_suffix_rows, arrow.rs:985-1019:
//                                      y_i = RN(b_i - sum_k<i L_ik y_k), one rounding
//   S_border -= Y^T RN(D^-1 Y)         contribution_product / subtract_contributions,
//                                      arrow.rs:1935-2074: one exact product per
//                                      leaf, subtracted in leaf order
//   border LDL^T                       factor_border, arrow.rs:3378-3404 -> ldl_rec
#include <omp.h>

#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "common.h"
#include "ldl.h"
#include "mpx.h"
#include "sampled.h"

namespace sdpx {
namespace {

using mpx::View;

void product(int limbs, const LdlOptions &opt, char ta, char tb, size_t m, size_t n, size_t k,
             const uint8_t *a, size_t lda, const uint8_t *b, size_t ldb, uint8_t *c, bool packed,
             SampledStats &st) {
    LdlStats ls;
    exact_product_public(limbs, opt, ta, tb, m, n, k, a, lda, b, ldb, c, packed, ls);
    st.gpu_products += ls.gpu_products;
    st.host_products += ls.host_products;
}

}  // namespace

bool sampled_arrow_factor(int limbs, const std::vector<SampledLeafIn> &leaves, size_t t,
                          uint8_t *border, const int8_t *border_signs, size_t nborder_signs,
                          const LdlOptions &opt, std::vector<SampledLeafOut> &out,
                          uint8_t *border_dinv, SampledStats &st) {
    const int N = limbs;
    const size_t vb = value_bytes(N);
    out.assign(leaves.size(), SampledLeafOut{});
    std::vector<uint8_t> half(vb);
    {
        // 1/2 exactly
        std::memset(half.data(), 0, vb);
        ((uint64_t *)half.data())[N - 1] = 1ULL << 63;
        *(int32_t *)(half.data() + 8 * N) = mpx::KIND_REG;
        *(int32_t *)(half.data() + 8 * N + 4) = 0;
    }
    double t0 = omp_get_wtime();
    for (size_t b = 0; b < leaves.size(); b++) {
        const SampledLeafIn &in = leaves[b];
        SampledLeafOut &o = out[b];
        const size_t h = in.h, K = in.k;
        // Schur build: V = Rinv Q (h x K), Gram = V^T V (packed upper, K x K)
        std::vector<uint8_t> V(h * K * vb), G(K * (K + 1) / 2 * vb);
        product(N, opt, 'N', 'N', h, K, h, in.rinv, h, in.q, h, V.data(), false, st);
        product(N, opt, 'T', 'N', K, K, h, V.data(), h, V.data(), h, G.data(), true, st);
        // leaf H (K x K, lower triangle), entry (p, q) as sampled/mod.rs:1973-1988
        o.l.assign(K * K * vb, 0);
        o.dinv.assign(K * vb, 0);
        for (size_t e = 0; e < K * K; e++) mpx::set_zero(&o.l[e * vb], N);
#pragma omp parallel num_threads(opt.threads)
        {
            std::vector<uint8_t> t1(vb), t2(vb), u1(vb), u(vb);
#pragma omp for schedule(dynamic, 4)
            for (int64_t qq = 0; qq < (int64_t)K; qq++) {
                const size_t q = (size_t)qq;
                for (size_t p = q; p < K; p++) {
                    const uint8_t *kpq = &G[(p * (p + 1) / 2 + q) * vb];  // (q, p), q <= p
                    const uint8_t *wp = in.w + p * vb, *wq = in.w + q * vb;
                    mpx::mul(View{N, wp}, View{N, wq}, t1.data());           // w_p w_q
                    mpx::mul(View{N, t1.data()}, View{N, half.data()}, t2.data());  // * 1/2
                    mpx::mul(View{N, kpq}, View{N, kpq}, u1.data());         // K K
                    mpx::add(View{N, u1.data()}, View{N, u1.data()}, u.data());  // + K K
                    mpx::mul(View{N, t2.data()}, View{N, u.data()}, &o.l[(p + q * K) * vb]);
                }
            }
        }
    }
    st.build_s += omp_get_wtime() - t0;
    t0 = omp_get_wtime();
    // leaves: factor, Y, contributions
    std::vector<uint8_t> tmp(vb);
    for (size_t b = 0; b < leaves.size(); b++) {
        const SampledLeafIn &in = leaves[b];
        SampledLeafOut &o = out[b];
        const size_t K = in.k;
        LdlStats ls;
        std::vector<int8_t> plus(1, 1);
        if (!ldl_factor(N, o.l.data(), K, K, plus.data(), 1, opt, o.dinv.data(), ls, false)) {
            st.failed_leaf = (int64_t)b;
            return false;
        }
        st.regularized += ls.regularized;
        std::vector<size_t> first = envelope(N, o.l.data(), K, K);
        // Y = L^-1 B by rows, one exact dot per entry (coupling B: K x t)
        o.y.assign(K * t * vb, 0);
        std::memcpy(o.y.data(), in.b, K * t * vb);
        std::vector<uint8_t> negl(K * K * vb);
        for (size_t e = 0; e < K * K; e++) {
            mpx::copy_value(&negl[e * vb], &o.l[e * vb], N);
            mpx::negate(&negl[e * vb], N);
        }
        std::vector<uint8_t> one(vb);
        std::memset(one.data(), 0, vb);
        ((uint64_t *)one.data())[N - 1] = 1ULL << 63;
        *(int32_t *)(one.data() + 8 * N) = mpx::KIND_REG;
        *(int32_t *)(one.data() + 8 * N + 4) = 1;
#pragma omp parallel num_threads(opt.threads)
        {
            mpx::DotAcc acc;
            std::vector<uint8_t> row((K + 1) * vb), res(vb);
#pragma omp for schedule(dynamic, 8)
            for (int64_t cc = 0; cc < (int64_t)t; cc++) {
                uint8_t *col = &o.y[(size_t)cc * K * vb];
                for (size_t i = 0; i < K; i++) {
                    const size_t lo = first[i];
                    // staged row: -L_ik for lo <= k < i, then 1
                    for (size_t k = lo; k < i; k++) mpx::copy_value(&row[(k - lo) * vb], &negl[(i + k * K) * vb], N);
                    mpx::copy_value(&row[(i - lo) * vb], one.data(), N);
                    acc.dot(row.data(), vb, col + lo * vb, vb, (int64_t)(i - lo + 1), N, res.data());
                    mpx::copy_value(col + i * vb, res.data(), N);
                }
            }
        }
        // Z = RN(Y D^-1) entrywise, contribution = upper(Y^T Z) exact
        std::vector<uint8_t> z(K * t * vb), prod(t * (t + 1) / 2 * vb);
#pragma omp parallel for num_threads(opt.threads) schedule(static)
        for (int64_t e = 0; e < (int64_t)(K * t); e++)
            mpx::mul(View{N, &o.y[e * vb]}, View{N, &o.dinv[(e % K) * vb]}, &z[e * vb]);
        product(N, opt, 'T', 'N', t, t, K, o.y.data(), K, z.data(), K, prod.data(), true, st);
        // border (upper triangle, t x t, ld t) -= contribution, in leaf order
#pragma omp parallel num_threads(opt.threads)
        {
            std::vector<uint8_t> r(vb);
#pragma omp for schedule(dynamic, 16)
            for (int64_t jj = 0; jj < (int64_t)t; jj++) {
                const size_t j = (size_t)jj;
                for (size_t i = 0; i <= j; i++) {
                    uint8_t *s = border + (i + j * t) * vb;
                    mpx::sub(View{N, s}, View{N, &prod[(j * (j + 1) / 2 + i) * vb]}, r.data());
                    mpx::copy_value(s, r.data(), N);
                }
            }
        }
    }
    st.leaves_s += omp_get_wtime() - t0;
    t0 = omp_get_wtime();
    // border: mirror the upper triangle and factor recursively
    for (size_t j = 0; j < t; j++)
        for (size_t i = j + 1; i < t; i++)
            mpx::copy_value(border + (i + j * t) * vb, border + (j + i * t) * vb, N);
    LdlStats ls;
    bool ok = ldl_factor(N, border, t, t, border_signs, nborder_signs, opt, border_dinv, ls, t >= 256);
    st.gpu_products += ls.gpu_products;
    st.host_products += ls.host_products;
    st.regularized += ls.regularized;
    st.border_s += omp_get_wtime() - t0;
    return ok;
}

}  // namespace sdpx
