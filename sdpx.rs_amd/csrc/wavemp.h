// Wave-distributed multiple-precision arithmetic for the latency-bound
// chains of the panel kernels (one factor column's pivot, one triangular
// solve step): a value is spread over a whole wave, lane k holding 32-bit
// mantissa word k (lanes >= W hold zero) with kind and exponent uniform
// over the wave.  Word-parallel steps replace the serial word loops of the
// register routines in mpdev.h:
//   - carries and borrows resolve in one step from two ballots (the lanes
//     that generate and propagate a carry), the chain itself formed by one
//     64-bit addition over the lane masks;
//   - word shifts are lane permutations (ds_bpermute), products are one
//     32x32-bit multiply per lane per word of the left operand;
//   - the rounding point, round bit and sticky bit come from ballots and
//     lane reads.
// Every routine gives the bits of its register counterpart (add_signed,
// mul, recip, the exact-window rounding): RN of the exact result, ties to
// even, the same overflow and underflow rules.  All 64 lanes must be active
// and the operands wave-uniform in kind and exponent.
#pragma once
#include "mpdev.h"

namespace sdpx {
namespace wmp {

using mpd::EMAX;
using mpd::EMIN;
using mpd::KIND_INF;
using mpd::KIND_REG;
using mpd::KIND_ZERO;
using mpd::Val;

template <int W>
struct WV {
    uint32_t w;  // this lane's mantissa word (0 at lanes >= W)
    int32_t kind, e;
};

__device__ __forceinline__ int lane() { return (int)__lane_id(); }
__device__ __forceinline__ uint32_t rl(uint32_t v, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// lane k receives lane k - 1's value, lane 0 receives 0
__device__ __forceinline__ uint32_t up1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}
// lane k receives lane src's value (0 when src lies outside the wave)
__device__ __forceinline__ uint32_t from_lane(uint32_t v, int src) {
    const uint32_t r = (uint32_t)__builtin_amdgcn_ds_bpermute((src & 63) << 2, (int)v);
    return src >= 0 && src < 64 ? r : 0u;
}

template <int W>
__device__ __forceinline__ bool is_zero(const WV<W> &v) {
    return v.kind == KIND_ZERO || v.kind == -KIND_ZERO;
}
template <int W>
__device__ __forceinline__ bool is_regular(const WV<W> &v) {
    return v.kind == KIND_REG || v.kind == -KIND_REG;
}
template <int W>
__device__ __forceinline__ WV<W> zero() {
    return WV<W>{0u, KIND_ZERO, 0};
}

// A record (W words, kind, exponent) in one coalesced load / store.
template <int W>
__device__ __forceinline__ WV<W> load(const uint32_t *rec) {
    const int k = lane();
    const uint32_t v = k < W + 2 ? rec[k] : 0u;
    return WV<W>{k < W ? v : 0u, (int32_t)rl(v, W), (int32_t)rl(v, W + 1)};
}
// The same in two halves, so a load can be issued a step ahead: the lane's
// word of the record (load_word), then the value (from_word, which waits).
template <int W>
__device__ __forceinline__ uint32_t load_word(const uint32_t *rec) {
    const int k = lane();
    return k < W + 2 ? rec[k] : 0u;
}
template <int W>
__device__ __forceinline__ WV<W> from_word(uint32_t v) {
    return WV<W>{lane() < W ? v : 0u, (int32_t)rl(v, W), (int32_t)rl(v, W + 1)};
}
template <int W>
__device__ __forceinline__ void store(uint32_t *rec, const WV<W> &x) {
    const int k = lane();
    if (k < W + 2) rec[k] = k < W ? x.w : (k == W ? (uint32_t)x.kind : (uint32_t)x.e);
}
// Register form (every lane the same value) and back.
template <int W>
__device__ __forceinline__ Val<W> to_val(const WV<W> &x) {
    Val<W> v;
#pragma unroll
    for (int i = 0; i < W; i++) v.m[i] = rl(x.w, i);
    v.kind = x.kind;
    v.e = x.e;
    return v;
}
template <int W>
__device__ __forceinline__ WV<W> from_val(const Val<W> &v) {
    const int k = lane();
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < W; i++) w = k == i ? v.m[i] : w;
    return WV<W>{w, (int32_t)__builtin_amdgcn_readfirstlane(v.kind), (int32_t)__builtin_amdgcn_readfirstlane(v.e)};
}

// Carry lookahead over lanes: lane k holds word s and sends carry bit cout
// (0 or 1) to lane k + 1.  A lane generates a carry when its incoming bit
// wraps it and propagates one when it ends all ones; the carries into every
// lane are then ((G | P) + G) ^ P over the 64-bit lane masks.
__device__ __forceinline__ uint32_t resolve_carry(uint32_t s, uint32_t cout) {
    const uint32_t a = up1(cout);
    const uint32_t x = s + a;
    const uint64_t G = ballot(a != 0u && x == 0u), P = ballot(x == 0xFFFFFFFFu);
    const uint64_t C = ((G | P) + G) ^ P;
    return x + (uint32_t)((C >> lane()) & 1u);
}
// The same for borrows: lane k holds word s and owes bout (0 or 1) to lane
// k + 1.
__device__ __forceinline__ uint32_t resolve_borrow(uint32_t s, uint32_t bout) {
    const uint32_t b = up1(bout);
    const uint32_t x = s - b;
    const uint64_t G = ballot(b != 0u && s == 0u), P = ballot(x == 0u);
    const uint64_t C = ((G | P) + G) ^ P;
    return x - (uint32_t)((C >> lane()) & 1u);
}

// RN((-1)^neg * sum_k z_k 2^(32 k) * 2^scale) to W words (z: this lane's
// word; any lane may be non-zero): round_words of the same buffer.
template <int W>
__device__ __forceinline__ WV<W> round(bool neg, uint32_t z, int64_t scale) {
    constexpr int P = 32 * W;
    const int k = lane();
    const uint64_t nz = ballot(z != 0u);
    if (nz == 0) return zero<W>();
    const int top = 63 - __builtin_clzll(nz);
    const uint32_t tw = rl(z, top);
    int shift = 32 * top + (32 - __builtin_clz(tw)) - P;
    const int q = shift >> 5, r = shift & 31;  // floor division
    const uint32_t lo = from_lane(z, q + k), hi = from_lane(z, q + k + 1);
    uint32_t m = k < W ? __builtin_amdgcn_alignbit(hi, lo, (uint32_t)r) : 0u;
    bool roundb = false, sticky = false;
    if (shift > 0) {
        const int rb = shift - 1, qr = rb >> 5, br = rb & 31;
        const uint32_t wr = rl(z, qr);
        roundb = (wr >> br) & 1u;
        sticky = (br && (wr & ((1u << br) - 1u)) != 0u) || (nz & ((1ull << qr) - 1ull)) != 0;
    }
    if (roundb && (sticky || (rl(m, 0) & 1u))) {
        const uint64_t ones = ballot(k < W && m == 0xFFFFFFFFu);
        const uint64_t A = (ones + 1) ^ ones;  // lanes 0 .. first not-all-ones word
        if ((A >> W) & 1u) {                   // every word was all ones
            m = k == W - 1 ? 0x80000000u : 0u;
            shift++;
        } else {
            m += (uint32_t)((A >> k) & 1u);
        }
    }
    const int64_t e = scale + shift + P;
    if (e > EMAX) return WV<W>{0u, neg ? -KIND_INF : KIND_INF, 0};
    if (e < EMIN) return zero<W>();
    return WV<W>{m, neg ? -KIND_REG : KIND_REG, (int32_t)e};
}

// RN(a + s*b), s = -1 when negate_b: the exact sum over the wave (the
// larger-exponent operand shifted up by the exponent gap, at most P + 1
// bits), then one rounding.  add_signed's result.
template <int W>
__device__ __forceinline__ WV<W> add(const WV<W> &a, const WV<W> &b, bool negate_b) {
    if (is_zero(b)) return a;
    if (is_zero(a)) {
        WV<W> r = b;
        if (negate_b && is_regular(r)) r.kind = -r.kind;
        return r;
    }
    constexpr int P = 32 * W;
    const int k = lane();
    const bool na = a.kind < 0, nb = (b.kind < 0) != negate_b;
    const bool swap = b.e > a.e;
    const uint32_t xw = swap ? b.w : a.w, yw = swap ? a.w : b.w;
    const int32_t xe = swap ? b.e : a.e, ye = swap ? a.e : b.e;
    const bool nx = swap ? nb : na, ny = swap ? na : nb;
    const int64_t gap = (int64_t)xe - ye;
    if (gap >= P + 2) return WV<W>{xw, nx ? -KIND_REG : KIND_REG, xe};
    const int wq = (int)(gap >> 5), bq = (int)(gap & 31);
    const uint32_t c0 = from_lane(xw, k - wq), c1 = from_lane(xw, k - wq - 1);
    const uint32_t X = bq ? __builtin_amdgcn_alignbit(c0, c1, (uint32_t)(32 - bq)) : c0;  // Mx << gap
    const uint32_t Y = yw;
    const int64_t scale = (int64_t)ye - P;
    if (nx == ny) {
        const uint32_t s = X + Y;
        return round<W>(nx, resolve_carry(s, s < X ? 1u : 0u), scale);
    }
    const uint64_t dm = ballot(X != Y);
    if (dm == 0) return zero<W>();
    const int t = 63 - __builtin_clzll(dm);
    const bool xbig = rl(X, t) > rl(Y, t);
    const uint32_t big = xbig ? X : Y, small = xbig ? Y : X;
    return round<W>(xbig ? nx : ny, resolve_borrow(big - small, big < small ? 1u : 0u), scale);
}

// The exact product's words, lane k holding word k of Ma * Mb (2W words):
// column sums of the 32x32-bit products (the right operand shifted up one
// lane per word of the left), then the carries.
template <int W>
__device__ __forceinline__ uint32_t product_words(const WV<W> &a, const WV<W> &b) {
    uint64_t lo = 0, hi = 0;
    uint32_t bs = b.w;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t p = (uint64_t)rl(a.w, i) * bs;
        lo += (uint32_t)p;
        hi += (uint32_t)(p >> 32);
        bs = up1(bs);
    }
    // column k: lo_k + hi_{k-1} (each sum of at most W words: below 2^37)
    const uint64_t hin = ((uint64_t)up1((uint32_t)(hi >> 32)) << 32) | up1((uint32_t)hi);
    const uint64_t c = lo + hin;
    const uint32_t x = (uint32_t)c, ci = up1((uint32_t)(c >> 32));
    const uint32_t s = x + ci;
    return resolve_carry(s, s < x ? 1u : 0u);
}

// RN(a * b): mul's result.
template <int W>
__device__ __forceinline__ WV<W> mul(const WV<W> &a, const WV<W> &b) {
    if (is_zero(a) || is_zero(b)) return zero<W>();
    constexpr int P = 32 * W;
    return round<W>((a.kind < 0) != (b.kind < 0), product_words(a, b), (int64_t)a.e + b.e - 2 * P);
}

// RN(1 / d) for a normalized d: Q = floor(2^(2P+1) / D) by base-2^32
// schoolbook division with the remainder window on lanes 0..W, shifted up
// one lane per quotient digit; each digit's multiply-subtract is one
// product per lane and one borrow resolution.  Then (Q << 1 | sticky)
// rounded once: recip's result.
//
// Quotient digits: q = floor(R / V) (R < 2^32 V) is estimated in binary64 as
// R3 / V2 from R's top three words and V's top two.  Truncating both moves
// the quotient by at most (q + 2) / V2 < 2^-30, and forming R3, V2, 1 / V2
// and the product rounds six times at most (relative 2^-53 each), so the
// estimate lies within 2^-18 of R / V: when its fraction is at least 2^-16
// from an integer, its floor is q.  Otherwise the digit comes from Knuth's
// estimate (algorithm D, TAOCP vol. 2, 4.3.1: two-word estimate, the
// v_{n-2} test, add back when the remainder goes negative).
template <int W>
__device__ __forceinline__ WV<W> recip(const WV<W> &d) {
    if (is_zero(d)) return WV<W>{0u, d.kind < 0 ? -KIND_INF : KIND_INF, 0};
    constexpr int P = 32 * W, QW = W + 2;
    const int k = lane();
    const uint32_t v = d.w;
    const uint64_t vt = rl(v, W - 1), vs = rl(v, W - 2);
    const double rcp = 1.0 / (double)vt;
    const double rv2 = 1.0 / fma((double)vt, 4294967296.0, (double)vs);  // 1 / V2
    // Q < 2^(P+2) has W + 1 digits (its digit W + 1 is 0): the first window
    // is numerator words u[W .. 2W] with u[2W] = 2
    uint32_t R = k == W ? 2u : 0u;
    uint32_t Q = 0;
#pragma unroll 1
    for (int j = QW - 2; j >= 0; j--) {
        if (j < QW - 2) R = up1(R);  // next window: R * 2^32 (the numerator's lower words are 0)
        const uint32_t r2 = rl(R, W), r1 = rl(R, W - 1), r0 = rl(R, W - 2);
        const double est = fma(fma((double)r2, 4294967296.0, (double)r1), 4294967296.0, (double)r0) * rv2;
        const double fl = floor(est), fr = est - fl;
        uint64_t qh;
        if (fr >= 0x1p-16 && fr <= 1.0 - 0x1p-16) {
            qh = (uint32_t)fl;  // exact (see above); below 2^32 since R < 2^32 V
        } else {
            const uint64_t num = ((uint64_t)r2 << 32) | r1;
            uint64_t rh;
            qh = mpd::qdigit(num, vt, rcp, rh);
            while (qh >> 32 || qh * vs > ((rh << 32) | r0)) {
                qh--;
                rh += vt;
                if (rh >> 32) break;
            }
        }
        // R -= qh * V over lanes 0..W
        const uint64_t p = (uint64_t)(uint32_t)qh * v;
        const uint32_t plo = (uint32_t)p, phi = up1((uint32_t)(p >> 32));
        const int64_t t = (int64_t)R - plo - phi;  // in (-2^33, 2^32)
        const uint32_t x = (uint32_t)t, bo = (uint32_t)(-(t >> 32));
        const uint32_t bi = up1(bo);
        const uint32_t s = x - bi;
        uint32_t Rn = resolve_borrow(s, x < bi ? 1u : 0u);
        if (rl(Rn, W + 1) != 0u) {  // negative: qh was one too large, add V back
            qh--;
            Rn = k <= W ? Rn : 0u;
            const uint32_t s2 = Rn + v;
            Rn = resolve_carry(s2, s2 < Rn ? 1u : 0u);
        }
        R = k <= W ? Rn : 0u;
        Q = k == j ? (uint32_t)qh : Q;
    }
    const bool sticky = ballot(R != 0u) != 0;
    const uint32_t mag = (Q << 1) | (up1(Q) >> 31) | (k == 0 && sticky ? 1u : 0u);
    return round<W>(d.kind < 0, mag, -(int64_t)P - 2 - d.e);
}

// An exact dot's window: lane s holds slot s's signed 64-bit total
// (weight 2^(32 s) at scale base; lanes >= SLOTS hold 0, the top slot keeps
// the sign).  The carries resolve in fixed steps: each slot below the top
// keeps its low word and sends the rest (|T| < 2^62, so under 2^30) up one
// lane; the words then carry at most +1 or -1 into the next lane, and those
// resolve by lookahead over the lanes below the top (the +1 carries, then
// the -1 borrows), whatever leaves the top word going into the top slot's
// signed total.  A negative total is negated across lanes, then one
// rounding: RN of the window's exact value, as the serial carry, negation
// and round_words give it.
template <int W, int SLOTS>
__device__ __forceinline__ WV<W> round_window(int64_t T, int64_t base) {
    static_assert(SLOTS <= 64, "one slot a lane");
    const int k = lane();
    const bool below_top = k < SLOTS - 1;
    // multi-bit carries, one lane up
    const int32_t c1 = below_top ? (int32_t)(T >> 32) : 0;
    int64_t U = (below_top ? (int64_t)(uint32_t)(uint64_t)T : T) + (int32_t)up1((uint32_t)c1);
    // below the top: U in (-2^31, 2^32 + 2^31), a word and a carry of -1, 0 or +1
    const int32_t c2 = below_top ? (int32_t)(U >> 32) : 0;
    const uint32_t ap = up1(c2 == 1 ? 1u : 0u), an = up1(c2 == -1 ? 1u : 0u);
    uint32_t x = (uint32_t)(uint64_t)U + (below_top ? ap : 0u);
    const uint64_t G = ballot(below_top && ap != 0u && x == 0u), P = ballot(below_top && x == 0xFFFFFFFFu);
    const uint64_t C = ((G | P) + G) ^ P;  // +1 carries into each lane
    const uint32_t cp = (uint32_t)((C >> k) & 1u);
    x += below_top ? cp : 0u;
    const uint32_t x0 = x;
    x -= below_top ? an : 0u;
    const uint64_t G2 = ballot(below_top && an != 0u && x0 == 0u), P2 = ballot(below_top && x == 0u);
    const uint64_t B = ((G2 | P2) + G2) ^ P2;  // -1 borrows into each lane
    const uint32_t bn = (uint32_t)((B >> k) & 1u);
    x -= below_top ? bn : 0u;
    if (k == SLOTS - 1) U += (int64_t)ap + cp - an - bn;  // the top slot takes what leaves the words below
    const int64_t ttop = (int64_t)(((uint64_t)rl((uint32_t)((uint64_t)U >> 32), SLOTS - 1) << 32) |
                                   rl((uint32_t)(uint64_t)U, SLOTS - 1));
    const bool neg = ttop < 0;
    uint32_t w = below_top ? x : (k == SLOTS - 1 ? (uint32_t)(uint64_t)U : 0u);
    if (neg) {
        const uint64_t nzm = ballot(w != 0u);
        const int f = __builtin_ctzll(nzm);
        w = k < f ? 0u : (k == f ? 0u - w : ~w);
        if (k >= SLOTS) w = 0u;
    }
    return round<W>(neg, w, base);
}

// T += s * (Ma * Mb) << off (off >= 0 bits above slot 0; the product's
// words must land below slot SLOTS - 1): lane s's slot total takes word s
// of the shifted exact product.
template <int W, int SLOTS>
__device__ __forceinline__ int64_t add_product(int64_t T, const WV<W> &a, const WV<W> &b, int64_t off, bool negate) {
    const int k = lane();
    const uint32_t z = product_words(a, b);
    const int wo = (int)(off >> 5), bo = (int)(off & 31);
    const uint32_t c0 = from_lane(z, k - wo), c1 = from_lane(z, k - wo - 1);
    const uint32_t word = bo ? __builtin_amdgcn_alignbit(c0, c1, (uint32_t)(32 - bo)) : c0;
    const bool neg = ((a.kind < 0) != (b.kind < 0)) != negate;
    return k < SLOTS ? (neg ? T - (int64_t)word : T + (int64_t)word) : T;
}

// T += s * z << off for a magnitude spread one word a lane (z: this lane's
// word; off >= 0 bits above slot 0, landing below slot SLOTS - 1).
template <int SLOTS>
__device__ __forceinline__ int64_t add_words(int64_t T, uint32_t z, int64_t off, bool negate) {
    const int k = lane();
    const int wo = (int)(off >> 5), bo = (int)(off & 31);
    const uint32_t c0 = from_lane(z, k - wo), c1 = from_lane(z, k - wo - 1);
    const uint32_t word = bo ? __builtin_amdgcn_alignbit(c0, c1, (uint32_t)(32 - bo)) : c0;
    return k < SLOTS ? (negate ? T - (int64_t)word : T + (int64_t)word) : T;
}

// RN(c - a * b) with one rounding: fms's result.  c's words and the exact
// product's 2W words go into one exact window at their bit offsets above
// the lower of their two scales (the span stays below 3P + 2 bits, SLOTS =
// 3W + 4), then round_window.  Two far cases: a product below a quarter of
// c's last unit leaves c (RN(c - t) = c for 0 < |t| < ulp(c) / 4, also at a
// power of two); a c below the product's last unit by two bits is replaced
// by a same-signed unit one slot below the product (both lie strictly
// between the same multiples of the product's last unit, so they round
// alike).
template <int W>
__device__ __forceinline__ WV<W> fms(const WV<W> &c, const WV<W> &a, const WV<W> &b) {
    if (is_zero(a) || is_zero(b)) return is_zero(c) ? zero<W>() : c;
    constexpr int P = 32 * W, SLOTS = 3 * W + 4;
    static_assert(SLOTS <= 64, "one slot a lane");
    const int64_t pe = (int64_t)a.e + b.e - 2 * P;  // scale of the product's word 0
    if (is_zero(c)) return round<W>((a.kind < 0) == (b.kind < 0), product_words(a, b), pe);
    const int64_t ce = (int64_t)c.e - P;  // scale of c's word 0: c's last unit
    if (pe + 2 * P <= ce - 2) return c;
    int64_t T, base;
    if (ce + P <= pe - 2) {
        base = pe - 32;
        T = lane() == 0 ? (c.kind < 0 ? -1 : 1) : 0;
        T = add_product<W, SLOTS>(T, a, b, 32, true);
    } else {
        base = ce < pe ? ce : pe;
        T = add_words<SLOTS>(0, c.w, ce - base, c.kind < 0);
        T = add_product<W, SLOTS>(T, a, b, pe - base, true);
    }
    return round_window<W, SLOTS>(T, base);
}

// Sign test of a - b for regular or zero operands: -1, 0 or 1 (the sign of
// add(a, b, true) without rounding it).
template <int W>
__device__ __forceinline__ int compare(const WV<W> &a, const WV<W> &b) {
    const bool za = is_zero(a), zb = is_zero(b);
    if (za && zb) return 0;
    const int sa = za ? 0 : (a.kind < 0 ? -1 : 1), sb = zb ? 0 : (b.kind < 0 ? -1 : 1);
    if (sa != sb) return sa > sb ? 1 : -1;
    // same sign, both non-zero: compare magnitudes (normalized mantissas)
    int mag = 0;
    if (a.e != b.e) {
        mag = a.e > b.e ? 1 : -1;
    } else {
        const uint64_t dm = ballot(a.w != b.w);
        if (dm) {
            const int t = 63 - __builtin_clzll(dm);
            mag = rl(a.w, t) > rl(b.w, t) ? 1 : -1;
        }
    }
    return sa > 0 ? mag : -mag;
}

}  // namespace wmp
}  // namespace sdpx
