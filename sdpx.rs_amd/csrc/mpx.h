// Host multiple-precision arithmetic on the MpFloat<N> interchange layout:
// correctly rounded (nearest-even) products, sums, reciprocals and exact dot
// products, as the panel steps of the LDL^T need them.
//
// Semantics follow the reference's scalar type (crates/arithmetic/src/
// lib.rs:333-340; MPFR round-to-nearest-even on every operation) and its
// exact dot (crates/arithmetic/src/exactdot.rs:1-13: exact accumulation,
// one rounding, from_scaled_integer 1316-1391).  The code is this
// library's own: sums align at most 2P + 2 bits (a farther operand only
// decides the sticky bit), reciprocals use long division.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace sdpx {
namespace mpx {

typedef unsigned __int128 u128;
constexpr int KIND_NAN = 0, KIND_INF = 1, KIND_ZERO = 2, KIND_REG = 3;
constexpr int64_t EMAX = (1LL << 30) - 1, EMIN = 1 - (1LL << 30);
constexpr int MAXN = 16;

// A value: limbs[n] (LSB first), kind, exponent; stride 8n + 8 bytes.
struct View {
    int n;
    const uint8_t *p;
    const uint64_t *limbs() const { return (const uint64_t *)p; }
    int32_t kind() const { return *(const int32_t *)(p + 8 * n); }
    int32_t exponent() const { return *(const int32_t *)(p + 8 * n + 4); }
    bool zero() const { return kind() == KIND_ZERO || kind() == -KIND_ZERO; }
    bool regular() const { return kind() == KIND_REG || kind() == -KIND_REG; }
    bool finite() const { return zero() || regular(); }
    bool negative() const { return kind() < 0; }
};

inline void set_zero(uint8_t *out, int n) {
    std::memset(out, 0, 8 * (size_t)n);
    *(int32_t *)(out + 8 * n) = KIND_ZERO;
    *(int32_t *)(out + 8 * n + 4) = 0;
}
inline void set_special(uint8_t *out, int n, int kind) {
    std::memset(out, 0, 8 * (size_t)n);
    *(int32_t *)(out + 8 * n) = kind;
    *(int32_t *)(out + 8 * n + 4) = 0;
}
inline void copy_value(uint8_t *out, const uint8_t *in, int n) { std::memcpy(out, in, 8 * (size_t)n + 8); }
inline void negate(uint8_t *v, int n) {
    int32_t *k = (int32_t *)(v + 8 * n);
    if (*k == KIND_REG || *k == -KIND_REG || *k == KIND_INF || *k == -KIND_INF) *k = -*k;
}

inline int bitlen(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// Round (-1)^neg * mag * 2^scale (mag: len limbs, LSB first) once to 64n bits.
inline void round_scaled(bool neg, const uint64_t *mag, int len, int64_t scale, int n, uint8_t *out) {
    while (len > 0 && mag[len - 1] == 0) len--;
    if (len == 0) {
        set_zero(out, n);
        return;
    }
    const int64_t bits = (int64_t)(len - 1) * 64 + bitlen(mag[len - 1]);
    const int64_t p = 64 * (int64_t)n;
    int64_t shift = bits - p;
    uint64_t m[MAXN];
    for (int l = 0; l < n; l++) {
        int64_t start = shift + 64 * (int64_t)l;
        int64_t src = start >= 0 ? start / 64 : -((-start + 63) / 64);
        int off = (int)(start - src * 64);
        uint64_t v = 0;
        if (src >= 0 && src < len) v |= mag[src] >> off;
        if (off && src + 1 >= 0 && src + 1 < len) v |= mag[src + 1] << (64 - off);
        m[l] = v;
    }
    if (shift > 0) {
        const int64_t rb = shift - 1;
        const bool round = (mag[rb / 64] >> (rb % 64)) & 1;
        bool sticky = false;
        for (int64_t w = 0; w < rb / 64 && !sticky; w++) sticky = mag[w] != 0;
        if (!sticky && (rb % 64)) sticky = (mag[rb / 64] & ((1ULL << (rb % 64)) - 1)) != 0;
        if (round && (sticky || (m[0] & 1))) {
            bool carry = true;
            for (int l = 0; l < n && carry; l++) carry = ++m[l] == 0;
            if (carry) {
                m[n - 1] = 1ULL << 63;
                shift++;
            }
        }
    }
    const int64_t e = scale + shift + p;
    if (e > EMAX) {
        set_special(out, n, neg ? -KIND_INF : KIND_INF);
        return;
    }
    if (e < EMIN) {
        set_zero(out, n);
        return;
    }
    std::memcpy(out, m, 8 * (size_t)n);
    *(int32_t *)(out + 8 * n) = neg ? -KIND_REG : KIND_REG;
    *(int32_t *)(out + 8 * n + 4) = (int32_t)e;
}

// RN(a * b)
inline void mul(const View &a, const View &b, uint8_t *out) {
    const int n = a.n;
    if (!a.finite() || !b.finite()) {
        set_special(out, n, KIND_NAN);  // not produced on the admitted path
        return;
    }
    if (a.zero() || b.zero()) {
        set_zero(out, n);
        return;
    }
    uint64_t prod[2 * MAXN] = {0};
    const uint64_t *x = a.limbs(), *y = b.limbs();
    for (int i = 0; i < n; i++) {
        u128 carry = 0;
        for (int j = 0; j < n; j++) {
            u128 t = (u128)x[i] * y[j] + prod[i + j] + carry;
            prod[i + j] = (uint64_t)t;
            carry = t >> 64;
        }
        prod[i + n] = (uint64_t)carry;
    }
    round_scaled(a.negative() != b.negative(), prod, 2 * n,
                 (int64_t)a.exponent() + b.exponent() - 128 * (int64_t)n, n, out);
}

// RN(a + b); sub: RN(a - b).
inline void add_signed(const View &a, bool nega, const View &b, bool negb, uint8_t *out) {
    const int n = a.n;
    if (a.zero() && b.zero()) {
        set_zero(out, n);
        return;
    }
    if (b.zero()) {
        copy_value(out, a.p, n);
        if (nega != a.negative()) negate(out, n);
        return;
    }
    if (a.zero()) {
        copy_value(out, b.p, n);
        if (negb != b.negative()) negate(out, n);
        return;
    }
    // order so that x has the larger exponent
    const View *x = &a, *y = &b;
    bool nx = nega, ny = negb;
    if (b.exponent() > a.exponent()) {
        std::swap(x, y);
        std::swap(nx, ny);
    }
    const int64_t p = 64 * (int64_t)n;
    const int64_t gap = (int64_t)x->exponent() - y->exponent();
    // |y| < 2^(e_y) <= 2^(e_x - P - 2): below a quarter ulp of x's smallest
    // neighbour spacing, so RN(x + y) = x (no tie can arise).
    if (gap >= p + 2) {
        copy_value(out, x->p, n);
        if (nx != x->negative()) negate(out, n);
        return;
    }
    // exact: X = Mx << gap, Y = My, both at scale e_y - P
    const int len = 2 * n + 2;
    uint64_t X[2 * MAXN + 2] = {0}, Y[2 * MAXN + 2] = {0};
    const int w = (int)(gap / 64), off = (int)(gap % 64);
    const uint64_t *mx = x->limbs();
    for (int i = 0; i < n; i++) {
        X[i + w] |= off ? (mx[i] << off) : mx[i];
        if (off) X[i + w + 1] |= mx[i] >> (64 - off);
    }
    std::memcpy(Y, y->limbs(), 8 * (size_t)n);
    if (nx == ny) {
        u128 c = 0;
        for (int i = 0; i < len; i++) {
            u128 t = (u128)X[i] + Y[i] + c;
            X[i] = (uint64_t)t;
            c = t >> 64;
        }
        round_scaled(nx, X, len, (int64_t)y->exponent() - p, n, out);
        return;
    }
    // different signs: subtract the smaller magnitude
    int cmp = 0;
    for (int i = len - 1; i >= 0 && !cmp; i--)
        if (X[i] != Y[i]) cmp = X[i] > Y[i] ? 1 : -1;
    if (cmp == 0) {
        set_zero(out, n);
        return;
    }
    uint64_t *big = cmp > 0 ? X : Y, *small = cmp > 0 ? Y : X;
    uint64_t borrow = 0;
    for (int i = 0; i < len; i++) {
        u128 d = (u128)big[i] - small[i] - borrow;
        big[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) ? 1 : 0;
    }
    round_scaled(cmp > 0 ? nx : ny, big, len, (int64_t)y->exponent() - p, n, out);
}
inline void add(const View &a, const View &b, uint8_t *out) {
    add_signed(a, a.negative(), b, b.negative(), out);
}
inline void sub(const View &a, const View &b, uint8_t *out) {
    add_signed(a, a.negative(), b, !b.negative(), out);
}

// RN(1 / d) by restoring long division: Q = floor(2^(2P+1) / M), sticky = rem != 0.
inline void recip(const View &d, uint8_t *out) {
    const int n = d.n;
    if (d.zero()) {
        set_special(out, n, d.negative() ? -KIND_INF : KIND_INF);
        return;
    }
    const uint64_t *m = d.limbs();
    const int64_t p = 64 * (int64_t)n;
    // quotient of 2^(2P+1) / M has P+1 or P+2 bits; compute P+2 quotient bits
    const int qbits = (int)(p + 2);
    std::vector<uint64_t> rem(n + 1, 0), q((qbits + 63) / 64 + 1, 0);
    // numerator 2^(2P+1): process bits from the top; rem starts at 0
    const int64_t total = 2 * p + 2;  // bit positions 2P+1 .. 0
    for (int64_t bit = total - 1; bit >= 0; bit--) {
        // rem = rem*2 + numerator bit
        uint64_t carry = (bit == 2 * p + 1) ? 1 : 0;
        for (int i = 0; i <= n; i++) {
            uint64_t v = rem[i];
            rem[i] = (v << 1) | carry;
            carry = v >> 63;
        }
        // compare rem with M
        int c = 0;
        if (rem[n]) c = 1;
        for (int i = n - 1; i >= 0 && !c; i--)
            if (rem[i] != m[i]) c = rem[i] > m[i] ? 1 : -1;
        if (c >= 0) {
            uint64_t borrow = 0;
            for (int i = 0; i <= n; i++) {
                uint64_t mi = i < n ? m[i] : 0;
                u128 dd = (u128)rem[i] - mi - borrow;
                rem[i] = (uint64_t)dd;
                borrow = (uint64_t)(dd >> 64) ? 1 : 0;
            }
            if (bit < (int64_t)q.size() * 64) q[bit / 64] |= 1ULL << (bit % 64);
        }
    }
    bool sticky = false;
    for (auto v : rem) sticky |= v != 0;
    // value = (Q + frac) * 2^-(2P+1) / 2^(e - P) ... 1/d = 2^(P - e) / M
    // = (2^(2P+1)/M) * 2^(-P-1-e); append a sticky bit below Q.
    std::vector<uint64_t> mag(q.size() + 1, 0);
    // mag = Q << 1 | sticky
    uint64_t c = sticky ? 1 : 0;
    for (size_t i = 0; i < q.size(); i++) {
        mag[i] = (q[i] << 1) | c;
        c = q[i] >> 63;
    }
    mag[q.size()] = c;
    round_scaled(d.negative(), mag.data(), (int)mag.size(), -p - 1 - d.exponent() - 1, n, out);
}

// RN(a / b) by restoring long division: Q = floor(Ma 2^(P+1) / Mb) has P + 1
// or P + 2 bits, sticky = rem != 0.  a, b not NaN; b zero gives a signed
// infinity, a zero a signed zero (0/0 and inf/inf are not admitted).
inline void div(const View &a, const View &b, uint8_t *out) {
    const int n = a.n;
    const bool neg = a.negative() != b.negative();
    if (b.zero() || !a.finite()) {
        set_special(out, n, neg ? -KIND_INF : KIND_INF);
        return;
    }
    if (a.zero() || !b.finite()) {
        set_zero(out, n);
        if (neg) *(int32_t *)(out + 8 * n) = -KIND_ZERO;
        return;
    }
    const uint64_t *ma = a.limbs(), *mb = b.limbs();
    const int64_t p = 64 * (int64_t)n;
    std::vector<uint64_t> rem(n + 1, 0), q((p + 2 + 63) / 64 + 1, 0);
    // numerator Ma 2^(P+1): bit positions 2P .. 0, Ma's bits on top
    for (int64_t bit = 2 * p; bit >= 0; bit--) {
        const int64_t src = bit - (p + 1);
        uint64_t carry = src >= 0 ? (ma[src / 64] >> (src % 64)) & 1 : 0;
        for (int i = 0; i <= n; i++) {
            uint64_t v = rem[i];
            rem[i] = (v << 1) | carry;
            carry = v >> 63;
        }
        int c = rem[n] ? 1 : 0;
        for (int i = n - 1; i >= 0 && !c; i--)
            if (rem[i] != mb[i]) c = rem[i] > mb[i] ? 1 : -1;
        if (c >= 0) {
            uint64_t borrow = 0;
            for (int i = 0; i <= n; i++) {
                const uint64_t mi = i < n ? mb[i] : 0;
                const u128 dd = (u128)rem[i] - mi - borrow;
                rem[i] = (uint64_t)dd;
                borrow = (uint64_t)(dd >> 64) ? 1 : 0;
            }
            q[bit / 64] |= 1ULL << (bit % 64);
        }
    }
    bool sticky = false;
    for (auto v : rem) sticky |= v != 0;
    // a / b = Q 2^(-(P+1) + ea - eb) + (below Q's last bit): value = (2Q + sticky) 2^(ea - eb - P - 2)
    std::vector<uint64_t> mag(q.size() + 1, 0);
    uint64_t c = sticky ? 1 : 0;
    for (size_t i = 0; i < q.size(); i++) {
        mag[i] = (q[i] << 1) | c;
        c = q[i] >> 63;
    }
    mag[q.size()] = c;
    round_scaled(neg, mag.data(), (int)mag.size(), (int64_t)a.exponent() - b.exponent() - p - 2, n, out);
}

// Sign of |a| - |b| for values that are not NaN (zeros of either sign are
// equal and smallest, infinities largest).
inline int cmp_abs(const View &a, const View &b) {
    const int ca = a.zero() ? 0 : (a.regular() ? 1 : 2), cb = b.zero() ? 0 : (b.regular() ? 1 : 2);
    if (ca != cb) return ca < cb ? -1 : 1;
    if (ca != 1) return 0;
    if (a.exponent() != b.exponent()) return a.exponent() < b.exponent() ? -1 : 1;
    for (int i = a.n - 1; i >= 0; i--)
        if (a.limbs()[i] != b.limbs()[i]) return a.limbs()[i] < b.limbs()[i] ? -1 : 1;
    return 0;
}

// Sign of a - b for values that are not NaN (+0 == -0).
inline int cmp(const View &a, const View &b) {
    const int sa = a.zero() ? 0 : (a.negative() ? -1 : 1), sb = b.zero() ? 0 : (b.negative() ? -1 : 1);
    if (sa != sb) return sa < sb ? -1 : 1;
    return sa >= 0 ? cmp_abs(a, b) : -cmp_abs(a, b);
}

// Exact dot of k pairs (x_i, y_i) rounded once.  Non-finite terms are not
// admitted on this path (the caller's values are finite).
struct DotAcc {
    std::vector<uint64_t> acc;  // two's complement accumulator
    int n = 0;
    void dot(const uint8_t *x, size_t xs, const uint8_t *y, size_t ys, int64_t k, int nn,
             uint8_t *out, const uint8_t *sub_from = nullptr);
};

inline void DotAcc::dot(const uint8_t *x, size_t xs, const uint8_t *y, size_t ys, int64_t k, int nn,
                        uint8_t *out, const uint8_t *sub_from) {
    n = nn;
    const int64_t p = 64 * (int64_t)n;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t i = 0; i < k; i++) {
        View a{n, x + i * xs}, b{n, y + i * ys};
        if (a.zero() || b.zero()) continue;
        int64_t s = (int64_t)a.exponent() + b.exponent() - 2 * p;
        lo = std::min(lo, s);
        hi = std::max(hi, s);
    }
    View c{n, sub_from};
    const bool with_c = sub_from != nullptr;  // result c - sum (c may be zero)
    const bool c_nz = with_c && !c.zero();
    if (c_nz) {
        int64_t s = (int64_t)c.exponent() - p;
        lo = std::min(lo, s);
        hi = std::max(hi, s - p);  // treat as a 2P-bit term at scale s - P
    }
    if (lo == INT64_MAX) {
        set_zero(out, n);
        return;
    }
    const int64_t width = hi - lo + 2 * p + 64 + 64;
    const int len = (int)(width / 64 + 2);
    acc.assign(len, 0);
    uint64_t prod[2 * MAXN + 1];
    auto add_shift = [&](const uint64_t *v, int vl, int64_t shift, bool neg) {
        const int64_t w = shift / 64;
        const int off = (int)(shift % 64);
        if (!neg) {
            u128 carry = 0;
            int64_t j = 0;
            for (; j <= vl; j++) {
                uint64_t cur = j < vl ? v[j] : 0, prev = j ? v[j - 1] : 0;
                uint64_t s = off ? ((cur << off) | (prev >> (64 - off))) : cur;
                u128 t = (u128)acc[w + j] + s + carry;
                acc[w + j] = (uint64_t)t;
                carry = t >> 64;
            }
            for (int64_t t = w + j; carry && t < len; t++) {
                u128 u = (u128)acc[t] + carry;
                acc[t] = (uint64_t)u;
                carry = u >> 64;
            }
        } else {
            uint64_t borrow = 0;
            int64_t j = 0;
            for (; j <= vl; j++) {
                uint64_t cur = j < vl ? v[j] : 0, prev = j ? v[j - 1] : 0;
                uint64_t s = off ? ((cur << off) | (prev >> (64 - off))) : cur;
                u128 d = (u128)acc[w + j] - s - borrow;
                acc[w + j] = (uint64_t)d;
                borrow = (uint64_t)(d >> 64) ? 1 : 0;
            }
            for (int64_t t = w + j; borrow && t < len; t++) {
                u128 d = (u128)acc[t] - borrow;
                acc[t] = (uint64_t)d;
                borrow = (uint64_t)(d >> 64) ? 1 : 0;
            }
        }
    };
    for (int64_t i = 0; i < k; i++) {
        View a{n, x + i * xs}, b{n, y + i * ys};
        if (a.zero() || b.zero()) continue;
        const uint64_t *ma = a.limbs(), *mb = b.limbs();
        std::memset(prod, 0, 8 * (size_t)(2 * n));
        for (int u = 0; u < n; u++) {
            u128 carry = 0;
            for (int v = 0; v < n; v++) {
                u128 t = (u128)ma[u] * mb[v] + prod[u + v] + carry;
                prod[u + v] = (uint64_t)t;
                carry = t >> 64;
            }
            prod[u + n] = (uint64_t)carry;
        }
        int64_t s = (int64_t)a.exponent() + b.exponent() - 2 * p;
        add_shift(prod, 2 * n, s - lo, a.negative() != b.negative());
    }
    if (with_c) {
        // subtract the exact dot from c: acc = c - sum  ->  negate acc, add c
        for (auto &v : acc) v = ~v;
        bool carry = true;
        for (auto &v : acc) {
            if (!carry) break;
            carry = ++v == 0;
        }
        if (c_nz) add_shift(c.limbs(), n, (int64_t)c.exponent() - p - lo, c.negative());
    }
    const bool neg = (acc[len - 1] >> 63) != 0;
    if (neg) {
        for (auto &v : acc) v = ~v;
        bool carry = true;
        for (auto &v : acc) {
            if (!carry) break;
            carry = ++v == 0;
        }
    }
    round_scaled(neg, acc.data(), len, lo, n, out);
}

}  // namespace mpx
}  // namespace sdpx
