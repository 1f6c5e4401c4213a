// Multiple-precision dense LDL^T with exact products on the GPU.
//
// The recursion, its rounding sequence and the panel steps follow the
// reference's dense leaf factor (crates/solver/src/solver/kkt/ldl/arrow.rs):
//   ldl_dots            156-253  entrywise exact dots (n <= 128 or leaves < 256)
//   ldl_rec             255-415  envelope cuts, halving, W21 = A21 L11^-T,
//                                 L21 = RN(W21 D1^-1), A22 -= RN(W21 L21^T)
//   solve_unit_lower_t  417-490  recursive halves, exact products between,
//                                 rows by exact dots below 128 columns
//   DenseLeaf::pivot    618-634  non-finite / regularization / zero pivots
// so the factor is bitwise the reference's.  Every exact product (the
// trailing update and the solve's coupling products) runs as one RNS product
// on INT8 matrix cores (sdpx_gpu_rns_gemm semantics); small products and
// declined ones take host exact dots, which give the same values.
// Host panels run on OpenMP threads (rows are independent).
#include <omp.h>

#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "common.h"
#include "ldl.h"
#include "mpx.h"
#include "rns_gemm.h"
#include "sampled.h"

namespace sdpx {
namespace {

using mpx::View;

constexpr size_t REC_BASE = 128;

struct Ctx {
    int n;        // limbs
    size_t vb;    // bytes per value
    const LdlOptions *opt;
    LdlStats *stats;
    uint8_t *at(uint8_t *base, size_t idx) const { return base + idx * vb; }
    const uint8_t *at(const uint8_t *base, size_t idx) const { return base + idx * vb; }
};

// The zero value and one in this precision.
struct Consts {
    std::vector<uint8_t> zero, one, minus_one;
    explicit Consts(int n) : zero(8 * n + 8), one(8 * n + 8), minus_one(8 * n + 8) {
        mpx::set_zero(zero.data(), n);
        std::memset(one.data(), 0, 8 * n);
        ((uint64_t *)one.data())[n - 1] = 1ULL << 63;
        *(int32_t *)(one.data() + 8 * n) = mpx::KIND_REG;
        *(int32_t *)(one.data() + 8 * n + 4) = 1;
        minus_one = one;
        *(int32_t *)(minus_one.data() + 8 * n) = -mpx::KIND_REG;
    }
};

// ---------------------------------------------------------------------------
// exact products
// ---------------------------------------------------------------------------
// c (m x n, ldc = m) = op(a) op(b), each entry the exact dot rounded once.
void exact_product(const Ctx &cx, char ta, char tb, size_t m, size_t n, size_t k, const uint8_t *a,
                   size_t lda, const uint8_t *b, size_t ldb, uint8_t *c, bool upper_packed) {
    if (m == 0 || n == 0) return;
    const size_t outs = upper_packed ? n * (n + 1) / 2 : m * n;
    if (k == 0) {
        for (size_t e = 0; e < outs; e++) mpx::set_zero(cx.at(c, e), cx.n);
        return;
    }
    if (cx.opt->gpu_products && (double)m * n * k >= cx.opt->gpu_min_work) {
        int32_t r = sdpx_gpu_rns_gemm((uint32_t)cx.n, (uint8_t)ta, (uint8_t)tb, m, n, k, a, lda, b,
                                      ldb, upper_packed ? 1 : 0, c, outs);
        if (r == SDPX_GPU_OK) {
            cx.stats->gpu_products++;
            cx.stats->gpu_mp_macs += (double)k * outs;
            return;
        }
        if (r != SDPX_GPU_DECLINED) throw Error(r, "exact product on the GPU failed");
        cx.stats->declined++;
    }
    cx.stats->host_products++;
    const bool tan = ta == 'N', tbn = tb == 'N';
    const size_t sa = tan ? lda * cx.vb : cx.vb, sb = tbn ? cx.vb : ldb * cx.vb;
    const int64_t total = (int64_t)outs;
#pragma omp parallel num_threads(cx.opt->threads)
    {
        mpx::DotAcc acc;
#pragma omp for schedule(dynamic, 16)
        for (int64_t e = 0; e < total; e++) {
            size_t i, j;
            if (upper_packed) {
                j = (size_t)((std::sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
                while (j * (j + 1) / 2 > (size_t)e) j--;
                while ((j + 1) * (j + 2) / 2 <= (size_t)e) j++;
                i = e - j * (j + 1) / 2;
            } else {
                i = e % m;
                j = e / m;
            }
            const uint8_t *x = a + (tan ? i : i * lda) * cx.vb;
            const uint8_t *y = b + (tbn ? j * ldb : j) * cx.vb;
            acc.dot(x, sa, y, sb, (int64_t)k, cx.n, cx.at(c, e));
        }
    }
}

// ---------------------------------------------------------------------------
// panel steps
// ---------------------------------------------------------------------------
using Sign = std::function<int(size_t)>;

// DenseLeaf::pivot: finite check, regularization, zero check.
bool pivot(const Ctx &cx, uint8_t *d, int sign, const std::string &) {
    View v{cx.n, d};
    if (!v.finite()) return false;
    if (cx.opt->reg) {
        // d * sign < eps  (sign = +-1: the product is exact)
        std::vector<uint8_t> ds(cx.vb);
        mpx::copy_value(ds.data(), d, cx.n);
        if (sign < 0) mpx::negate(ds.data(), cx.n);
        View dv{cx.n, ds.data()}, ev{cx.n, cx.opt->reg_eps};
        std::vector<uint8_t> diff(cx.vb);
        mpx::sub(dv, ev, diff.data());  // sign of (d*sign - eps), exact in sign
        View df{cx.n, diff.data()};
        if (!df.zero() && df.negative()) {
            mpx::copy_value(d, cx.opt->reg_delta, cx.n);
            if (sign < 0) mpx::negate(d, cx.n);
            cx.stats->regularized++;
        }
    }
    View out{cx.n, d};
    return !out.zero();
}

inline size_t tri(size_t i) { return i * (i + 1) / 2; }

// ldl_dots (arrow.rs:156-253): left-looking, entry (i, j) one exact dot of
// rows of L and W = L D (rounded entrywise); the factor is written into the
// lower triangle of l with a unit diagonal, D^-1 into dinv.
bool ldl_dots(const Ctx &cx, uint8_t *l, size_t ld, size_t n, const size_t *first, const Sign &sign,
              uint8_t *dinv) {
    const int N = cx.n;
    const size_t vb = cx.vb;
    std::vector<uint8_t> r(tri(n) * vb), w(tri(n) * vb), d(n * vb);
    Consts k1(N);
    mpx::DotAcc acc;
    std::vector<uint8_t> t(vb), t2(vb);
    auto A = [&](size_t i, size_t j) { return cx.at(l, i + j * ld); };
    // entry (i, j): RN(RN(a_ij - RN(dot(r_i[lo..j], w_j[lo..j]))) * dinv_j)
    auto entry = [&](mpx::DotAcc &ac, size_t i, size_t j, uint8_t *out, uint8_t *tmp,
                     uint8_t *diff) {
        if (first[i] > j) {
            mpx::set_zero(out, N);
            return;
        }
        const size_t lo = std::max(first[i], first[j]);
        ac.dot(&r[(tri(i) + lo) * vb], vb, &w[(tri(j) + lo) * vb], vb, (int64_t)(j - lo), N, tmp);
        mpx::sub(View{N, A(i, j)}, View{N, tmp}, diff);
        mpx::mul(View{N, diff}, View{N, cx.at(dinv, j)}, out);
    };
    for (size_t j = 0; j < n; j++) {
        const size_t fj = first[j];
        uint8_t *dj = &d[j * vb];
        acc.dot(&r[(tri(j) + fj) * vb], vb, &w[(tri(j) + fj) * vb], vb, (int64_t)(j - fj), N, t.data());
        mpx::sub(View{N, A(j, j)}, View{N, t.data()}, dj);
        if (View{N, dj}.zero() && j > fj) {
            // exact difference rounded once (the rounded dot cancelled it)
            acc.dot(&r[(tri(j) + fj) * vb], vb, &w[(tri(j) + fj) * vb], vb, (int64_t)(j - fj), N,
                    dj, A(j, j));
        }
        if (!pivot(cx, dj, sign(j), "")) return false;
        mpx::recip(View{N, dj}, cx.at(dinv, j));
        const int64_t rows = (int64_t)(n - j - 1);
#pragma omp parallel num_threads(rows > 64 ? cx.opt->threads : 1)
        {
            mpx::DotAcc ac;
            std::vector<uint8_t> tmp(vb), diff(vb);
#pragma omp for schedule(static)
            for (int64_t u = 0; u < rows; u++) {
                const size_t i = j + 1 + (size_t)u;
                uint8_t *rij = &r[(tri(i) + j) * vb];
                entry(ac, i, j, rij, tmp.data(), diff.data());
                mpx::mul(View{N, rij}, View{N, dj}, &w[(tri(i) + j) * vb]);
            }
        }
    }
    for (size_t j = 0; j < n; j++) {
        mpx::copy_value(A(j, j), k1.one.data(), N);
        for (size_t i = j + 1; i < n; i++) mpx::copy_value(A(i, j), &r[(tri(i) + j) * vb], N);
    }
    return true;
}

// X L^T = B (solve_unit_lower_t, arrow.rs:417-490): B is m x n (ldb), L unit
// lower n x n (ldl); low[i] is row i's first nonzero column (<= n).
void solve_unit_lower_t(const Ctx &cx, size_t m, size_t n, const uint8_t *L, size_t ldl, uint8_t *B,
                        size_t ldb, const std::vector<size_t> &low) {
    const int N = cx.n;
    const size_t vb = cx.vb;
    if (m == 0 || n == 0) return;
    if (n <= REC_BASE) {
        std::vector<uint8_t> lr(n * n * vb);
        for (size_t j = 0; j < n; j++)
            for (size_t k = 0; k < n; k++) mpx::copy_value(&lr[(j * n + k) * vb], cx.at(L, j + k * ldl), N);
#pragma omp parallel num_threads(cx.opt->threads)
        {
            mpx::DotAcc ac;
            std::vector<uint8_t> x(n * vb), t(vb), u(vb);
#pragma omp for schedule(dynamic, 4)
            for (int64_t ii = 0; ii < (int64_t)m; ii++) {
                const size_t i = (size_t)ii;
                for (size_t j = 0; j < n; j++) mpx::copy_value(&x[j * vb], cx.at(B, i + j * ldb), N);
                const size_t lo = low[i];
                for (size_t j = lo; j < n; j++) {
                    ac.dot(&x[lo * vb], vb, &lr[(j * n + lo) * vb], vb, (int64_t)(j - lo), N, t.data());
                    mpx::sub(View{N, &x[j * vb]}, View{N, t.data()}, u.data());
                    mpx::copy_value(&x[j * vb], u.data(), N);
                }
                for (size_t j = 0; j < n; j++) mpx::copy_value(cx.at(B, i + j * ldb), &x[j * vb], N);
            }
        }
        cx.stats->base_rows += m;
        return;
    }
    const size_t n1 = n / 2, n2 = n - n / 2;
    std::vector<size_t> low1(m), low2(m);
    for (size_t i = 0; i < m; i++) {
        low1[i] = std::min(low[i], n1);
        low2[i] = low[i] > n1 ? low[i] - n1 : 0;
    }
    solve_unit_lower_t(cx, m, n1, L, ldl, B, ldb, low1);
    uint8_t *B2 = cx.at(B, n1 * ldb);
    std::vector<uint8_t> p(m * n2 * vb);
    exact_product(cx, 'N', 'T', m, n2, n1, B, ldb, cx.at(L, n1), ldl, p.data(), false);
    {
        std::vector<uint8_t> t(vb);
        for (size_t j = 0; j < n2; j++)
            for (size_t i = 0; i < m; i++) {
                uint8_t *v = cx.at(B2, i + j * ldb);
                mpx::sub(View{N, v}, View{N, &p[(i + j * m) * vb]}, t.data());
                mpx::copy_value(v, t.data(), N);
            }
    }
    solve_unit_lower_t(cx, m, n2, cx.at(L, n1 + n1 * ldl), ldl, B2, ldb, low2);
}

bool ldl_rec(const Ctx &cx, uint8_t *l, size_t ld, size_t n, const size_t *first, const Sign &sign,
             uint8_t *dinv) {
    const int N = cx.n;
    const size_t vb = cx.vb;
    if (n <= REC_BASE) return ldl_dots(cx, l, ld, n, first, sign, dinv);
    // the envelope cut nearest the middle, if any
    size_t low = n;
    bool have = false;
    size_t cut = 0;
    for (size_t c = n - 1; c >= 1; c--) {
        low = std::min(low, first[c]);
        auto dist = [&](size_t x) { return x > n / 2 ? x - n / 2 : n / 2 - x; };
        if (low >= c && (!have || dist(c) < dist(cut))) {
            cut = c;
            have = true;
        }
    }
    if (have) {
        const size_t c = cut;
        std::vector<size_t> tail(n - c);
        for (size_t i = c; i < n; i++) tail[i - c] = first[i] - c;
        Sign sign2 = [&](size_t k) { return sign(c + k); };
        // the two sides are independent; either order gives the same bits
        bool ok1 = ldl_rec(cx, l, ld, c, first, sign, dinv);
        bool ok2 = ldl_rec(cx, cx.at(l, c + c * ld), ld, n - c, tail.data(), sign2, cx.at(dinv, c));
        return ok1 && ok2;
    }
    const size_t n1 = n / 2, n2 = n - n / 2;
    if (!ldl_rec(cx, l, ld, n1, first, sign, dinv)) return false;
    std::vector<uint8_t> l11(n1 * n1 * vb);
    for (size_t j = 0; j < n1; j++) std::memcpy(&l11[j * n1 * vb], cx.at(l, j * ld), n1 * vb);
    std::vector<size_t> lowv(n2);
    for (size_t i = 0; i < n2; i++) lowv[i] = std::min(first[n1 + i], n1);
    uint8_t *a21 = cx.at(l, n1);
    solve_unit_lower_t(cx, n2, n1, l11.data(), n1, a21, ld, lowv);
    l11.clear();
    l11.shrink_to_fit();
    // W21 before scaling; L21 = RN(W21 D1^-1) in place
    std::vector<uint8_t> w(n2 * n1 * vb);
    for (size_t j = 0; j < n1; j++) std::memcpy(&w[j * n2 * vb], cx.at(a21, j * ld), n2 * vb);
    {
        const int64_t cols = (int64_t)n1;
#pragma omp parallel for num_threads(cx.opt->threads) schedule(static)
        for (int64_t jj = 0; jj < cols; jj++) {
            std::vector<uint8_t> t(vb);
            const size_t j = (size_t)jj;
            for (size_t i = 0; i < n2; i++) {
                uint8_t *v = cx.at(a21, i + j * ld);
                mpx::mul(View{N, v}, View{N, cx.at(dinv, j)}, t.data());
                mpx::copy_value(v, t.data(), N);
            }
        }
    }
    // packed (a, b), a <= b, of W21 L21^T is the Schur term of A22(b, a)
    std::vector<uint8_t> packed(n2 * (n2 + 1) / 2 * vb);
    exact_product(cx, 'N', 'T', n2, n2, n1, w.data(), n2, a21, ld, packed.data(), true);
    w.clear();
    w.shrink_to_fit();
    uint8_t *a22 = cx.at(l, n1 + n1 * ld);
    {
        const int64_t cols = (int64_t)n2;
#pragma omp parallel for num_threads(cx.opt->threads) schedule(dynamic, 8)
        for (int64_t aa = 0; aa < cols; aa++) {
            std::vector<uint8_t> t(vb);
            const size_t a = (size_t)aa;
            for (size_t b = a; b < n2; b++) {
                uint8_t *v = cx.at(a22, b + a * ld);
                mpx::sub(View{N, v}, View{N, &packed[(b * (b + 1) / 2 + a) * vb]}, t.data());
                mpx::copy_value(v, t.data(), N);
            }
        }
    }
    packed.clear();
    packed.shrink_to_fit();
    std::vector<size_t> tail(n2);
    for (size_t i = 0; i < n2; i++) tail[i] = first[n1 + i] > n1 ? first[n1 + i] - n1 : 0;
    Sign sign2 = [&](size_t k) { return sign(n1 + k); };
    return ldl_rec(cx, a22, ld, n2, tail.data(), sign2, cx.at(dinv, n1));
}

}  // namespace

void exact_product_public(int limbs, const LdlOptions &opt, char ta, char tb, size_t m, size_t n,
                          size_t k, const uint8_t *a, size_t lda, const uint8_t *b, size_t ldb,
                          uint8_t *c, bool packed, LdlStats &stats) {
    Ctx cx{limbs, value_bytes(limbs), &opt, &stats};
    exact_product(cx, ta, tb, m, n, k, a, lda, b, ldb, c, packed);
}

// Row envelope: first nonzero column of each row of the lower triangle.
std::vector<size_t> envelope(int limbs, const uint8_t *l, size_t ld, size_t n) {
    const size_t vb = value_bytes(limbs);
    std::vector<size_t> first(n);
    for (size_t i = 0; i < n; i++) {
        size_t k = 0;
        while (k < i && mpx::View{limbs, l + (i + k * ld) * vb}.zero()) k++;
        first[i] = k;
    }
    return first;
}

bool ldl_factor(int limbs, uint8_t *l, size_t ld, size_t n, const int8_t *signs, size_t nsigns,
                const LdlOptions &opt, uint8_t *dinv, LdlStats &stats, bool recursive) {
    Ctx cx{limbs, value_bytes(limbs), &opt, &stats};
    std::vector<size_t> first = envelope(limbs, l, ld, n);
    Sign sign = [&](size_t k) { return nsigns == 0 ? 1 : (int)signs[nsigns == 1 ? 0 : k]; };
    if (recursive) {
        // DenseLeaf::factor_loaded (arrow.rs:819-846): a pivot that one
        // rounding per level cancels to an exact zero fails the recursive
        // factor; the input is restored and factored by exact dots, and the
        // failed attempt's regularization count is dropped.
        const size_t vb = cx.vb;
        std::vector<uint8_t> saved(n * n * vb);
        for (size_t j = 0; j < n; j++) std::memcpy(&saved[j * n * vb], l + j * ld * vb, n * vb);
        const int64_t reg0 = stats.regularized;
        if (ldl_rec(cx, l, ld, n, first.data(), sign, dinv)) return true;
        for (size_t j = 0; j < n; j++) std::memcpy(l + j * ld * vb, &saved[j * n * vb], n * vb);
        stats.regularized = reg0;
        stats.retried++;
    }
    return ldl_dots(cx, l, ld, n, first.data(), sign, dinv);
}

}  // namespace sdpx
