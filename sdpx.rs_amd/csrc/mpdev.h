// Multiple-precision arithmetic for GPU panel kernels, on the MpFloat<N>
// interchange layout held as 32-bit words: correctly rounded (nearest-even)
// products, sums, reciprocals and exact dot products.  Everything is
// __host__ __device__ so the host instantiation can be checked bit for bit
// against the host library on a machine without a GPU.
//
// Exact dots accumulate every partial product word into signed 64-bit slots
// (one per 32-bit position of a fixed window, strided so per-thread slices
// of LDS are bank-conflict free) and resolve carries once at the end, then
// round once - the contract of crates/arithmetic/src/exactdot.rs:1-13.  A
// dot whose exponent spread does not fit the window reports failure and the
// caller takes the host path for it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdpx {
namespace mpd {

// The larger operations are inlined into every kernel: an out-of-line call
// passes its Val operands through the stack (private scratch memory), and
// those round trips cost microseconds per call on the GPU.
#define SDPX_MP_OUTLINE __host__ __device__ inline __attribute__((always_inline))

constexpr int KIND_ZERO = 2, KIND_REG = 3, KIND_INF = 1, KIND_NAN = 0;
constexpr int64_t EMAX = (1LL << 30) - 1, EMIN = 1 - (1LL << 30);

// A value in registers: W = 2N mantissa words (LSB first), kind, exponent.
template <int W>
struct Val {
    uint32_t m[W];
    int32_t kind, e;
};

template <int W>
__host__ __device__ inline void load(const uint32_t *rec, Val<W> &v) {
#pragma unroll
    for (int i = 0; i < W; i++) v.m[i] = rec[i];
    v.kind = (int32_t)rec[W];
    v.e = (int32_t)rec[W + 1];
}
template <int W>
__host__ __device__ inline void store(uint32_t *rec, const Val<W> &v) {
#pragma unroll
    for (int i = 0; i < W; i++) rec[i] = v.m[i];
    rec[W] = (uint32_t)v.kind;
    rec[W + 1] = (uint32_t)v.e;
}
template <int W>
__host__ __device__ inline void set_zero(Val<W> &v) {
#pragma unroll
    for (int i = 0; i < W; i++) v.m[i] = 0;
    v.kind = KIND_ZERO;
    v.e = 0;
}
template <int W>
__host__ __device__ inline bool is_zero(const Val<W> &v) {
    return v.kind == KIND_ZERO || v.kind == -KIND_ZERO;
}
template <int W>
__host__ __device__ inline bool is_regular(const Val<W> &v) {
    return v.kind == KIND_REG || v.kind == -KIND_REG;
}
__host__ __device__ inline int bitlen32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return x ? 32 - __clz((int)x) : 0;
#else
    return x ? 32 - __builtin_clz(x) : 0;
#endif
}

// (hi:lo) >> sh, 0 <= sh < 32: the low word of the 64-bit funnel shift
__host__ __device__ inline uint32_t funnel(uint32_t hi, uint32_t lo, int sh) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)sh);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
#endif
}

// Word shifts of a register array by a run-time count, in log steps so every
// register index is static: t >>= q words (zero fill), t <<= q words.
template <int L>
__host__ __device__ inline void shr_words(uint32_t (&t)[L], int q) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1)
        if (s < 2 * L && (q & s)) {
#pragma unroll
            for (int i = 0; i < L; i++) t[i] = i + s < L ? t[i + s] : 0u;
        }
}
template <int L>
__host__ __device__ inline void shl_words(uint32_t (&t)[L], int q) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1)
        if (s < 2 * L && (q & s)) {
#pragma unroll
            for (int i = L - 1; i >= 0; i--) t[i] = i - s >= 0 ? t[i - s] : 0u;
        }
}

// Round (-1)^neg * mag * 2^scale once (mag: L words, LSB first) into out.
// Every register index is static (log-step word shifts), so mag may live in
// registers.
template <int W, int L>
__host__ __device__ inline void round_words(bool neg, const uint32_t (&mag)[L], int64_t scale,
                                            Val<W> &out) {
    int top = -1;
    uint32_t topw = 0;
#pragma unroll
    for (int i = 0; i < L; i++)
        if (mag[i]) {
            top = i;
            topw = mag[i];
        }
    if (top < 0) {
        set_zero(out);
        return;
    }
    const int P = 32 * W;
    const int bits = 32 * top + bitlen32(topw);
    int shift = bits - P;
    bool round = false, sticky = false;
    if (shift > 0) {
        const int rb = shift - 1;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const int lo = 32 * i;
            if (rb >= lo && rb < lo + 32) {
                const int off = rb - lo;
                round = (mag[i] >> off) & 1;
                if (off && (mag[i] & ((1u << off) - 1))) sticky = true;
            } else if (lo + 31 < rb && mag[i]) {
                sticky = true;
            }
        }
    }
    // mantissa word u = bits [shift + 32u, +32) of mag
    constexpr int T = L > W ? L + 1 : W + 1;
    uint32_t t[T];
#pragma unroll
    for (int i = 0; i < T; i++) t[i] = i < L ? mag[i] : 0u;
    if (shift >= 0) {
        shr_words<T>(t, shift >> 5);
        const int rs = shift & 31;
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = funnel(t[u + 1], t[u], rs);
    } else {
        const int ls = -shift;
        shl_words<T>(t, ls >> 5);
        const int rs = ls & 31;
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = rs ? funnel(t[u], u ? t[u - 1] : 0u, 32 - rs) : t[u];
    }
    if (round && (sticky || (out.m[0] & 1))) {
        uint32_t c = 1;
#pragma unroll
        for (int u = 0; u < W; u++) {
            const uint32_t v = out.m[u] + c;
            c = (c && v == 0) ? 1 : 0;
            out.m[u] = v;
        }
        if (c) {
            out.m[W - 1] = 0x80000000u;
            shift++;
        }
    }
    const int64_t e = scale + shift + P;
    if (e > EMAX) {
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = 0;
        out.kind = neg ? -KIND_INF : KIND_INF;
        out.e = 0;
        return;
    }
    if (e < EMIN) {
        set_zero(out);
        return;
    }
    out.kind = neg ? -KIND_REG : KIND_REG;
    out.e = (int32_t)e;
}

// round_words for a buffer whose top word is non-zero (mag[L - 1] != 0,
// L > W): the rounding point then lies in word B = L - 1 - W at the bit
// below the top word's length, so the mantissa is one funnel shift of the
// top W + 1 words and no run-time word shift is needed.  Same result as
// round_words on the same buffer.
template <int W, int L>
__host__ __device__ inline void round_norm(bool neg, const uint32_t (&mag)[L], int64_t scale,
                                           Val<W> &out) {
    static_assert(L > W, "the buffer must extend below the mantissa");
    constexpr int B = L - 1 - W;
    const int bl = bitlen32(mag[L - 1]);  // 1 .. 32
    const int P = 32 * W;
    int shift = 32 * B + bl;
    if (bl == 32) {
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = mag[B + 1 + u];
    } else {
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = funnel(mag[B + 1 + u], mag[B + u], bl);
    }
    const uint32_t rw = mag[B];
    const bool round = (rw >> (bl - 1)) & 1u;
    bool sticky = (rw & ((1u << (bl - 1)) - 1u)) != 0;
#pragma unroll
    for (int i = 0; i < B; i++) sticky |= mag[i] != 0;
    if (round && (sticky || (out.m[0] & 1))) {
        uint32_t c = 1;
#pragma unroll
        for (int u = 0; u < W; u++) {
            const uint32_t v = out.m[u] + c;
            c = (c && v == 0) ? 1 : 0;
            out.m[u] = v;
        }
        if (c) {
            out.m[W - 1] = 0x80000000u;
            shift++;
        }
    }
    const int64_t e = scale + shift + P;
    if (e > EMAX) {
#pragma unroll
        for (int u = 0; u < W; u++) out.m[u] = 0;
        out.kind = neg ? -KIND_INF : KIND_INF;
        out.e = 0;
        return;
    }
    if (e < EMIN) {
        set_zero(out);
        return;
    }
    out.kind = neg ? -KIND_REG : KIND_REG;
    out.e = (int32_t)e;
}

// RN(a * b)
template <int W>
SDPX_MP_OUTLINE void mul(const Val<W> &a, const Val<W> &b, Val<W> &out) {
    if (is_zero(a) || is_zero(b)) {
        set_zero(out);
        return;
    }
    uint32_t p[2 * W];
#pragma unroll
    for (int i = 0; i < 2 * W; i++) p[i] = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            const uint64_t t = (uint64_t)a.m[i] * b.m[j] + p[i + j] + carry;
            p[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        p[i + W] = (uint32_t)carry;
    }
    // normalized mantissas put the product's top bit in its top word
    if (p[2 * W - 1])
        round_norm<W, 2 * W>((a.kind < 0) != (b.kind < 0), p, (int64_t)a.e + b.e - 2 * 32 * (int64_t)W, out);
    else
        round_words<W, 2 * W>((a.kind < 0) != (b.kind < 0), p,
                              (int64_t)a.e + b.e - 2 * 32 * (int64_t)W, out);
}

// RN(a + s*b) for non-zero a, b: the exact sum over 2W + 2 words, then one
// rounding (any mantissas; the general path of add_signed).
template <int W>
__host__ __device__ inline void add_signed_exact(const Val<W> &a, const Val<W> &b, bool negate_b,
                                                 Val<W> &out) {
    const bool na = a.kind < 0, nb = (b.kind < 0) != negate_b;
    const bool swap = b.e > a.e;
    Val<W> xv, yv;
#pragma unroll
    for (int i = 0; i < W; i++) {
        xv.m[i] = swap ? b.m[i] : a.m[i];
        yv.m[i] = swap ? a.m[i] : b.m[i];
    }
    xv.e = swap ? b.e : a.e;
    yv.e = swap ? a.e : b.e;
    const bool nx = swap ? nb : na, ny = swap ? na : nb;
    const Val<W> *x = &xv, *y = &yv;
    const int64_t P = 32 * W;
    const int64_t gap = (int64_t)x->e - y->e;
    if (gap >= P + 2) {
        out = xv;
        out.kind = nx ? -KIND_REG : KIND_REG;
        return;
    }
    // X = Mx << gap (gap <= P + 1), Y = My; both at scale e_y - P
    constexpr int L = 2 * W + 2;
    uint32_t X[L], Y[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        X[i] = 0;
        Y[i] = i < W ? y->m[i] : 0;
    }
    const int wsh = (int)(gap >> 5), bsh = (int)(gap & 31);
    // X = Mx << gap: bit shift into W + 1 words, then a word shift
#pragma unroll
    for (int i = 0; i <= W; i++) {
        const uint32_t cur = i < W ? x->m[i] : 0u, prev = i ? x->m[i - 1] : 0u;
        X[i] = bsh ? funnel(cur, prev, 32 - bsh) : cur;
    }
    shl_words<L>(X, wsh);
    if (nx == ny) {
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const uint64_t t = (uint64_t)X[i] + Y[i] + c;
            X[i] = (uint32_t)t;
            c = t >> 32;
        }
        round_words<W, L>(nx, X, (int64_t)y->e - P, out);
        return;
    }
    int cmp = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; i--)
        if (cmp == 0 && X[i] != Y[i]) cmp = X[i] > Y[i] ? 1 : -1;
    if (cmp == 0) {
        set_zero(out);
        return;
    }
    uint32_t D[L];
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const uint64_t big = cmp > 0 ? X[i] : Y[i], small = cmp > 0 ? Y[i] : X[i];
        const uint64_t d = big - small - borrow;
        D[i] = (uint32_t)d;
        borrow = (d >> 63) & 1;
    }
    round_words<W, L>(cmp > 0 ? nx : ny, D, (int64_t)y->e - P, out);
}

// RN(a + s*b) with s = +1 (add) or -1 (sub).  With both mantissas
// normalized (top bit set) the sum is formed at x's scale over W + 3 words,
// x = the operand of larger exponent shifted up 64 bits and y shifted down
// by the exponent gap, the bits y loses folded into one sticky bit at the
// bottom: a lost bit needs a gap above 64, which leaves the result's top bit
// at least P + 62 bits above the bottom, so the sticky bit only breaks ties
// and the rounding is the exact sum's.  A difference that loses at most one
// bit at the top is rounded from the top one or two words (round_norm), a
// deeper cancellation (exponent gap 0 or 1, exact) by round_words.
template <int W>
SDPX_MP_OUTLINE void add_signed(const Val<W> &a, const Val<W> &b, bool negate_b,
                                Val<W> &out) {
    if (is_zero(b)) {
        out = a;
        return;
    }
    if (is_zero(a)) {
        out = b;
        if (negate_b && is_regular(out)) out.kind = -out.kind;
        return;
    }
    if (!(a.m[W - 1] >> 31) || !(b.m[W - 1] >> 31)) {
        add_signed_exact(a, b, negate_b, out);
        return;
    }
    const bool na = a.kind < 0, nb = (b.kind < 0) != negate_b;
    const bool swap = b.e > a.e;
    const Val<W> &x = swap ? b : a, &y = swap ? a : b;
    const bool nx = swap ? nb : na, ny = swap ? na : nb;
    constexpr int P = 32 * W;
    const int64_t gap = (int64_t)x.e - y.e;
    if (gap >= P + 2) {
        out = x;
        out.kind = nx ? -KIND_REG : KIND_REG;
        return;
    }
    // X = Mx << 64, Y = (My << 64) >> gap over W + 3 words (top word spare)
    constexpr int L = W + 3;
    uint32_t X[L], Y[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        X[i] = i >= 2 && i < W + 2 ? x.m[i - 2] : 0u;
        Y[i] = i >= 2 && i < W + 2 ? y.m[i - 2] : 0u;
    }
    const int wq = (int)(gap >> 5), bq = (int)(gap & 31);
    bool sticky = false;
#pragma unroll
    for (int i = 0; i < L; i++) {
        if (i < wq) sticky |= Y[i] != 0;
        if (i == wq && bq) sticky |= (Y[i] & ((1u << bq) - 1u)) != 0;
    }
    shr_words<L>(Y, wq);
    if (bq) {
#pragma unroll
        for (int i = 0; i < L; i++) Y[i] = funnel(i + 1 < L ? Y[i + 1] : 0u, Y[i], bq);
    }
    bool neg = nx;
    if (nx == ny) {
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const uint64_t t = (uint64_t)X[i] + Y[i] + c;
            X[i] = (uint32_t)t;
            c = t >> 32;
        }
    } else {
        if (gap == 0) {  // exact: the larger mantissa first
            int cmp = 0;
#pragma unroll
            for (int i = L - 1; i >= 0; i--)
                if (cmp == 0 && X[i] != Y[i]) cmp = X[i] > Y[i] ? 1 : -1;
            if (cmp == 0) {
                set_zero(out);
                return;
            }
            if (cmp < 0) {
#pragma unroll
                for (int i = 0; i < L; i++) {
                    const uint32_t t = X[i];
                    X[i] = Y[i];
                    Y[i] = t;
                }
                neg = ny;
            }
        }
        // X - Y - sticky (X > Y + 1 whenever a bit was lost)
        int64_t br = sticky ? 1 : 0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const int64_t t = (int64_t)X[i] - Y[i] - br;
            X[i] = (uint32_t)t;
            br = t < 0 ? 1 : 0;
        }
    }
    if (sticky) X[0] |= 1u;
    const int64_t scale = (int64_t)x.e - P - 64;
    if (X[L - 1]) {
        round_norm<W, L>(neg, X, scale, out);
    } else if (X[L - 2]) {
        uint32_t X2[L - 1];
#pragma unroll
        for (int i = 0; i < L - 1; i++) X2[i] = X[i];
        round_norm<W, L - 1>(neg, X2, scale, out);
    } else {
        round_words<W, L>(neg, X, scale, out);
    }
}

// RN(c - a*b) with one rounding (the fused multiply-add of the triangular
// sweeps).  The product is exact in 2W words and c is widened to 2W words;
// the two are then combined as in add_signed.  An operand more than
// 64W + 2 bits below the other is replaced by one bit at the bottom of the
// buffer: both values lie strictly inside the same unit of the buffer's last
// bit, more than two bits below the W-word rounding point, so they round
// alike.
template <int W>
SDPX_MP_OUTLINE void fms(const Val<W> &c, const Val<W> &a, const Val<W> &b, Val<W> &out) {
    if (is_zero(a) || is_zero(b)) {
        if (is_zero(c))
            set_zero(out);
        else
            out = c;
        return;
    }
    constexpr int M = 2 * W;
    uint32_t p[M];
#pragma unroll
    for (int i = 0; i < M; i++) p[i] = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            const uint64_t t = (uint64_t)a.m[i] * b.m[j] + p[i + j] + carry;
            p[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        p[i + W] = (uint32_t)carry;
    }
    const bool np = (a.kind < 0) == (b.kind < 0);  // sign of -a*b
    const int64_t PM = 32 * M;
    if (is_zero(c)) {
        round_words<W, M>(np, p, (int64_t)a.e + b.e - PM, out);
        return;
    }
    // both mantissas normalised to M words: value = m * 2^(e - 32M)
    int64_t ep = (int64_t)a.e + b.e;
    if (!(p[M - 1] >> 31)) {
#pragma unroll
        for (int i = M - 1; i >= 0; i--) p[i] = funnel(p[i], i ? p[i - 1] : 0u, 31);
        ep -= 1;
    }
    const bool nc = c.kind < 0;
    const bool swap = ep > (int64_t)c.e;  // x: the operand with the larger exponent
    uint32_t xm[M], ym[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
        const uint32_t cw = i >= W ? c.m[i - W] : 0u;
        xm[i] = swap ? p[i] : cw;
        ym[i] = swap ? cw : p[i];
    }
    const int64_t ex = swap ? ep : (int64_t)c.e, ey = swap ? (int64_t)c.e : ep;
    const bool nx = swap ? np : nc, ny = swap ? nc : np;
    constexpr int L = 2 * M + 2;
    int64_t gap = ex - ey;
    const bool far = gap >= PM + 2;
    if (far) gap = PM + 2;
    uint32_t X[L], Y[L];
#pragma unroll
    for (int i = 0; i < L; i++) {
        X[i] = 0;
        Y[i] = far ? (i == 0 ? 1u : 0u) : (i < M ? ym[i] : 0u);
    }
    const int wsh = (int)(gap >> 5), bsh = (int)(gap & 31);
#pragma unroll
    for (int i = 0; i <= M; i++) {
        const uint32_t cur = i < M ? xm[i] : 0u, prev = i ? xm[i - 1] : 0u;
        X[i] = bsh ? funnel(cur, prev, 32 - bsh) : cur;
    }
    shl_words<L>(X, wsh);
    const int64_t scale = ex - gap - PM;
    if (nx == ny) {
        uint64_t cy = 0;
#pragma unroll
        for (int i = 0; i < L; i++) {
            const uint64_t t = (uint64_t)X[i] + Y[i] + cy;
            X[i] = (uint32_t)t;
            cy = t >> 32;
        }
        round_words<W, L>(nx, X, scale, out);
        return;
    }
    int cmp = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; i--)
        if (cmp == 0 && X[i] != Y[i]) cmp = X[i] > Y[i] ? 1 : -1;
    if (cmp == 0) {
        set_zero(out);
        return;
    }
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const uint64_t big = cmp > 0 ? X[i] : Y[i], small = cmp > 0 ? Y[i] : X[i];
        const uint64_t d = big - small - borrow;
        X[i] = (uint32_t)d;
        borrow = (d >> 63) & 1;
    }
    round_words<W, L>(cmp > 0 ? nx : ny, X, scale, out);
}

// ---------------------------------------------------------------------------
// Exact dot with a slot window: slots[k * stride] for k < SLOTS, each a
// signed 64-bit sum of 32-bit partial words at position k (bit 32k above
// the window base).
// ---------------------------------------------------------------------------
template <int W, int SLOTS>
struct DotWindow {
    int64_t *slots;
    int stride;
    int64_t base;  // scale of slot 0's bit 0
    __host__ __device__ void clear() {
        for (int k = 0; k < SLOTS; k++) slots[k * stride] = 0;
    }
    // Add s * (a*b) where the product's scale (value = Ma*Mb*2^scale) is
    // scale >= base and the product fits the window.
    __host__ __device__ void add_product(const Val<W> &a, const Val<W> &b, bool negate) {
        uint32_t p[2 * W];
#pragma unroll
        for (int i = 0; i < 2 * W; i++) p[i] = 0;
#pragma unroll
        for (int i = 0; i < W; i++) {
            uint64_t carry = 0;
#pragma unroll
            for (int j = 0; j < W; j++) {
                const uint64_t t = (uint64_t)a.m[i] * b.m[j] + p[i + j] + carry;
                p[i + j] = (uint32_t)t;
                carry = t >> 32;
            }
            p[i + W] = (uint32_t)carry;
        }
        const bool neg = ((a.kind < 0) != (b.kind < 0)) != negate;
        const int64_t scale = (int64_t)a.e + b.e - 2 * 32 * (int64_t)W;
        add_words<2 * W>(p, scale, neg);
    }
    // Add s * (value) of one operand (its mantissa at scale e - P).
    __host__ __device__ void add_value(const Val<W> &a, bool negate) {
        const bool neg = (a.kind < 0) != negate;
        add_words<W>(a.m, (int64_t)a.e - 32 * (int64_t)W, neg);
    }
    template <int L>
    __host__ __device__ void add_words(const uint32_t (&p)[L], int64_t scale, bool neg) {
        const int64_t off = scale - base;
        const int wo = (int)(off >> 5), bo = (int)(off & 31);
        // shifted words p << bo occupy L + 1 slots from wo
        uint32_t prev = 0;
#pragma unroll
        for (int i = 0; i <= L; i++) {
            const uint32_t cur = i < L ? p[i] : 0;
            const uint32_t v = bo ? ((cur << bo) | (prev >> (32 - bo))) : cur;
            prev = cur;
            int64_t *s = slots + (int64_t)(wo + i) * stride;
            *s = neg ? *s - (int64_t)v : *s + (int64_t)v;
        }
    }
    // Resolve carries and round once into out.  Returns false when the
    // value does not fit (never for admitted windows).
    __host__ __device__ void finish(Val<W> &out) {
        // carry-resolve into 32-bit words (two's complement), kept in slots
        int64_t carry = 0;
        for (int k = 0; k < SLOTS; k++) {
            int64_t *s = slots + (int64_t)k * stride;
            const int64_t v = *s + carry;
            // floor division by 2^32 (arithmetic shift)
            carry = v >> 32;
            *s = (int64_t)(uint32_t)(v & 0xFFFFFFFFLL);
        }
        const bool neg = carry < 0;
        if (neg) {
            uint64_t c = 1;
            for (int k = 0; k < SLOTS; k++) {
                int64_t *s = slots + (int64_t)k * stride;
                const uint64_t v = (uint64_t)(~(uint32_t)*s) + c;
                *s = (int64_t)(uint32_t)v;
                c = v >> 32;
            }
        }
        // top word, then gather the top W + 2 words around it into registers
        int top = -1;
        for (int k = SLOTS - 1; k >= 0 && top < 0; k--)
            if (slots[(int64_t)k * stride] != 0) top = k;
        if (top < 0) {
            set_zero(out);
            return;
        }
        // words [top - W - 1, top]; anything below is folded into a sticky word
        constexpr int G = W + 2;
        uint32_t g[G + 1];
        const int lo = top - G + 1;
        bool sticky = false;
        for (int k = 0; k < lo && k < SLOTS; k++)
            if (slots[(int64_t)k * stride] != 0) sticky = true;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int k = lo + i;
            g[i + 1] = k >= 0 ? (uint32_t)slots[(int64_t)k * stride] : 0u;
        }
        g[0] = sticky ? 1u : 0u;  // one sticky bit below the gathered words
        round_words<W, G + 1>(neg, g, base + 32 * (int64_t)(lo - 1), out);
    }
};

// floor(num / vt) and its remainder for num < 2^64 and vt in [2^31, 2^32)
// (a schoolbook quotient-digit estimate).  The binary64 estimate
// num * RN(1/vt) is within 3.3 2^-53 relative of num / vt, so below 2^33 it
// is off by less than one: two below it is a lower bound, and at most four
// steps of r -= vt make the pair exact.  Replaces a 64-bit integer division
// (a long software sequence on the GPU) on the critical path of every
// reciprocal.
__host__ __device__ inline uint64_t qdigit(uint64_t num, uint64_t vt, double rcp, uint64_t &rem) {
    uint64_t q = (uint64_t)((double)num * rcp);
    q = q > 2 ? q - 2 : 0;
    uint64_t r = num - q * vt;
    while (r >= vt) {
        r -= vt;
        q++;
    }
    rem = r;
    return q;
}

// RN(1 / d): Q = floor(2^(2P+1) / D) for the P-bit mantissa D (top bit set)
// and the remainder's sticky bit, then one rounding of (Q << 1 | sticky).
// Q comes from schoolbook division in base 2^32 (Knuth, TAOCP vol. 2,
// 4.3.1, algorithm D: two-word quotient-digit estimate, at most two
// corrections, multiply-subtract, add back), W + 2 digit steps.
template <int W>
SDPX_MP_OUTLINE void recip(const Val<W> &d, Val<W> &out) {
    if (is_zero(d)) {
        set_zero(out);
        out.kind = d.kind < 0 ? -KIND_INF : KIND_INF;
        return;
    }
    const int P = 32 * W;
    constexpr int QW = W + 2;  // quotient digits
    // numerator 2^(2P+1): digit 2W holds 2; one spare top digit
    uint32_t u[2 * W + 2];
#pragma unroll
    for (int i = 0; i < 2 * W + 2; i++) u[i] = 0;
    u[2 * W] = 2;
    const uint32_t *v = d.m;
    const uint64_t vt = v[W - 1], vs = W >= 2 ? v[W - 2] : 0;
    const double rcp = 1.0 / (double)vt;
    uint32_t q[QW + 1];
    q[QW] = 0;
#pragma unroll
    for (int j = QW - 1; j >= 0; j--) {
        const uint64_t num = ((uint64_t)u[j + W] << 32) | u[j + W - 1];
        uint64_t rh, qh = qdigit(num, vt, rcp, rh);
        const uint64_t ulow = (W >= 2 && j + W - 2 >= 0) ? u[j + W - 2 >= 0 ? j + W - 2 : 0] : 0;
        while (qh >> 32 || qh * vs > ((rh << 32) | ulow)) {
            qh--;
            rh += vt;
            if (rh >> 32) break;
        }
        // u[j .. j+W] -= qh * v
        uint64_t mc = 0;
        int64_t br = 0;
#pragma unroll
        for (int i = 0; i < W; i++) {
            const uint64_t p = qh * v[i] + mc;
            mc = p >> 32;
            const int64_t t = (int64_t)u[j + i] - (int64_t)(uint32_t)p + br;
            u[j + i] = (uint32_t)t;
            br = t >> 32;
        }
        const int64_t t = (int64_t)u[j + W] - (int64_t)mc + br;
        u[j + W] = (uint32_t)t;
        if (t < 0) {  // qh was one too large: add v back
            qh--;
            uint64_t c = 0;
#pragma unroll
            for (int i = 0; i < W; i++) {
                const uint64_t s2 = (uint64_t)u[j + i] + v[i] + c;
                u[j + i] = (uint32_t)s2;
                c = s2 >> 32;
            }
            u[j + W] += (uint32_t)c;
        }
        q[j] = (uint32_t)qh;
    }
    bool sticky = false;
#pragma unroll
    for (int i = 0; i < W; i++) sticky |= u[i] != 0;
    // mag = (q << 1) | sticky
    uint32_t mag[QW + 1];
    uint32_t c = sticky ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < QW; i++) {
        mag[i] = (q[i] << 1) | c;
        c = q[i] >> 31;
    }
    mag[QW] = c;
    // Q lies in (2^(P+1), 2^(P+2)], so (Q << 1 | sticky) tops out in word W
    uint32_t mw[W + 1];
#pragma unroll
    for (int i = 0; i <= W; i++) mw[i] = mag[i];
    if (mw[W] && !mag[W + 1] && !mag[W + 2])
        round_norm<W, W + 1>(d.kind < 0, mw, -(int64_t)P - 2 - d.e, out);
    else
        round_words<W, QW + 1>(d.kind < 0, mag, -(int64_t)P - 2 - d.e, out);
}

// RN(x / d): Q = floor(Mx 2^(P+2) / Md) by the same base-2^32 schoolbook
// division (W + 2 quotient digits), the remainder's sticky bit, one rounding
// of (Q << 1 | sticky) 2^(ex - ed - P - 3).
template <int W>
SDPX_MP_OUTLINE void div(const Val<W> &x, const Val<W> &d, Val<W> &out) {
    if (is_zero(x)) {
        set_zero(out);
        return;
    }
    if (is_zero(d)) {
        set_zero(out);
        out.kind = (x.kind < 0) != (d.kind < 0) ? -KIND_INF : KIND_INF;
        return;
    }
    const int P = 32 * W;
    constexpr int QW = W + 2;
    // numerator Mx << (P + 2): digits W .. 2W hold Mx << 2
    uint32_t u[2 * W + 2];
#pragma unroll
    for (int i = 0; i < 2 * W + 2; i++) u[i] = 0;
#pragma unroll
    for (int i = 0; i <= W; i++) {
        const uint32_t cur = i < W ? x.m[i] : 0u, prev = i ? x.m[i - 1] : 0u;
        u[W + i] = funnel(cur, prev, 30);
    }
    const uint32_t *v = d.m;
    const uint64_t vt = v[W - 1], vs = W >= 2 ? v[W - 2] : 0;
    const double rcp = 1.0 / (double)vt;
    uint32_t q[QW + 1];
    q[QW] = 0;
#pragma unroll
    for (int j = QW - 1; j >= 0; j--) {
        const uint64_t num = ((uint64_t)u[j + W] << 32) | u[j + W - 1];
        uint64_t rh, qh = qdigit(num, vt, rcp, rh);
        const uint64_t ulow = u[j + W - 2 >= 0 ? j + W - 2 : 0];
        while (qh >> 32 || qh * vs > ((rh << 32) | (j + W - 2 >= 0 ? ulow : 0))) {
            qh--;
            rh += vt;
            if (rh >> 32) break;
        }
        uint64_t mc = 0;
        int64_t br = 0;
#pragma unroll
        for (int i = 0; i < W; i++) {
            const uint64_t pr = qh * v[i] + mc;
            mc = pr >> 32;
            const int64_t t = (int64_t)u[j + i] - (int64_t)(uint32_t)pr + br;
            u[j + i] = (uint32_t)t;
            br = t >> 32;
        }
        const int64_t t = (int64_t)u[j + W] - (int64_t)mc + br;
        u[j + W] = (uint32_t)t;
        if (t < 0) {
            qh--;
            uint64_t c = 0;
#pragma unroll
            for (int i = 0; i < W; i++) {
                const uint64_t s2 = (uint64_t)u[j + i] + v[i] + c;
                u[j + i] = (uint32_t)s2;
                c = s2 >> 32;
            }
            u[j + W] += (uint32_t)c;
        }
        q[j] = (uint32_t)qh;
    }
    bool sticky = false;
#pragma unroll
    for (int i = 0; i < W; i++) sticky |= u[i] != 0;
    uint32_t mag[QW + 1];
    uint32_t c = sticky ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < QW; i++) {
        mag[i] = (q[i] << 1) | c;
        c = q[i] >> 31;
    }
    mag[QW] = c;
    // Q lies in (2^(P+1), 2^(P+3)), so (Q << 1 | sticky) tops out in word W
    uint32_t mw[W + 1];
#pragma unroll
    for (int i = 0; i <= W; i++) mw[i] = mag[i];
    if (mw[W] && !mag[W + 1] && !mag[W + 2])
        round_norm<W, W + 1>((x.kind < 0) != (d.kind < 0), mw, (int64_t)x.e - d.e - P - 3, out);
    else
        round_words<W, QW + 1>((x.kind < 0) != (d.kind < 0), mag,
                               (int64_t)x.e - d.e - P - 3, out);
}

// RN(sqrt(a)) for a > 0: restoring integer square root, two radicand bits a
// step, of the mantissa padded to an even exponent and 2P + 4 or more bits
// (P + 3 root bits), the remainder's sticky bit, one rounding.
template <int W>
SDPX_MP_OUTLINE void sqrt(const Val<W> &a, Val<W> &out) {
    if (is_zero(a)) {
        set_zero(out);
        return;
    }
    const int P = 32 * W;
    // value = M 2^e; radicand R = M 2^ex with ex = (e odd) + 2k, bits >= 2P + 4
    int64_t e = (int64_t)a.e - P;
    int ex = 0;
    if (e & 1) {
        ex = 1;
        e -= 1;
    }
    const int k = (2 * P + 4 - (P + ex) + 1) / 2;
    ex += 2 * k;
    int rbits = P + ex;
    rbits += rbits & 1;
    constexpr int RW = W + 3;  // root and remainder words (P + 3 root bits)
    uint32_t rem[RW + 1], root[RW + 1];
#pragma unroll
    for (int i = 0; i <= RW; i++) rem[i] = root[i] = 0;
    for (int b = rbits - 2; b >= 0; b -= 2) {
        const int s1 = b + 1 - ex, s0 = b - ex;
        const uint32_t hi = s1 >= 0 && s1 < P ? (a.m[s1 >> 5] >> (s1 & 31)) & 1u : 0u;
        const uint32_t lo = s0 >= 0 && s0 < P ? (a.m[s0 >> 5] >> (s0 & 31)) & 1u : 0u;
        // rem = (rem << 2) | hi lo; trial = (root << 2) | 1
        uint32_t in = (hi << 1) | lo;
#pragma unroll
        for (int i = 0; i <= RW; i++) {
            const uint32_t vv = rem[i];
            rem[i] = (vv << 2) | in;
            in = vv >> 30;
        }
        uint32_t trial[RW + 1];
        uint32_t tin = 1;
#pragma unroll
        for (int i = 0; i <= RW; i++) {
            trial[i] = (root[i] << 2) | tin;
            tin = root[i] >> 30;
        }
        bool ge = true;
        bool decided = false;
#pragma unroll
        for (int i = RW; i >= 0; i--)
            if (!decided && rem[i] != trial[i]) {
                ge = rem[i] > trial[i];
                decided = true;
            }
        uint32_t rin = 0;
#pragma unroll
        for (int i = 0; i <= RW; i++) {
            const uint32_t vv = root[i];
            root[i] = (vv << 1) | rin;
            rin = vv >> 31;
        }
        if (ge) {
            int64_t br = 0;
#pragma unroll
            for (int i = 0; i <= RW; i++) {
                const int64_t t = (int64_t)rem[i] - trial[i] - br;
                rem[i] = (uint32_t)t;
                br = t < 0;
            }
            root[0] |= 1u;
        }
    }
    bool sticky = false;
#pragma unroll
    for (int i = 0; i <= RW; i++) sticky |= rem[i] != 0;
    uint32_t mag[RW + 2];
    uint32_t c = sticky ? 1u : 0u;
#pragma unroll
    for (int i = 0; i <= RW; i++) {
        mag[i] = (root[i] << 1) | c;
        c = root[i] >> 31;
    }
    mag[RW + 1] = c;
    round_words<W, RW + 2>(false, mag, e / 2 - k - 1, out);
}

}  // namespace mpd
}  // namespace sdpx
