// GPU panel kernels of the multiple-precision LDL^T and sampled arrow
// factor (arithmetic in mpdev.h).  Each kernel computes exactly the
// roundings of the corresponding host step (ldl.cpp / sampled.cpp), so
// moving a step to the GPU never changes a bit:
//   ew_sub        S = RN(S - P) over a packed upper triangle
//                 (subtract_contributions, arrow.rs:1975-2074; A22 update,
//                 arrow.rs:376-392)
//   ew_mul_diag   Z = RN(Y D) with D per row or column (contribution_product
//                 Z = D^-1 Y, arrow.rs:1940-1956; L21 = W21 D1^-1, 347-351)
//   forward_rows  Y = L^-1 B, y_i = RN(b_i - sum_k L_ik y_k) with one
//                 rounding (forward_suffix_rows, arrow.rs:1000-1019)
//   trsm_base     X L^T = B rows, x_j = RN(x_j - RN(dot)) (solve_unit_lower_t
//                 base case, arrow.rs:425-447)
//   ldl_dots      left-looking LDL^T of a block (arrow.rs:156-253), one
//                 workgroup per block
//   sampled_entries  H_pq = RN(RN(w_p w_q) RN(K_pq^2)) (entry,
//                 sampled/mod.rs:1973-1988, whose halving and doubling are exact)
// Exact dots use a slot window in LDS; a dot whose exponent spread exceeds
// the window raises a flag and the host recomputes the affected output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>

#include "common.h"
#include "mp_kernels.h"
#include "mpdev.h"
#include "wavemp.h"

namespace sdpx {

using namespace mpd;

template <int W>
__device__ __forceinline__ const uint32_t *rec_at(const uint32_t *base, int64_t idx) {
    return base + idx * (W + 2);
}
template <int W>
__device__ __forceinline__ uint32_t *rec_at(uint32_t *base, int64_t idx) {
    return base + idx * (W + 2);
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
// mode 0: S(i, j) (i <= j, upper, ld) -= P[j(j+1)/2 + i]
// mode 1: S(j, i) (lower mirror, b = j >= a = i) -= P[j(j+1)/2 + i]
template <int W>
__global__ void ew_sub_packed_kernel(uint32_t *S, int64_t ld, const uint32_t *P, int64_t n,
                                     int mode) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = n * (n + 1) / 2;
    if (e >= total) return;
    int64_t j = (int64_t)((::sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
    while (j * (j + 1) / 2 > e) j--;
    while ((j + 1) * (j + 2) / 2 <= e) j++;
    const int64_t i = e - j * (j + 1) / 2;
    const int64_t o = mode == 0 ? i + j * ld : j + i * ld;
    Val<W> s, p, r;
    load(rec_at<W>(S, o), s);
    load(rec_at<W>(P, e), p);
    add_signed(s, p, true, r);
    store(rec_at<W>(S, o), r);
}

// Z[i + j*ldz] = RN(Y[i + j*ldy] * D[per_row ? i : j]) for an m x n block
template <int W>
__global__ void ew_mul_diag_kernel(const uint32_t *Y, int64_t ldy, const uint32_t *D,
                                   uint32_t *Z, int64_t ldz, int64_t m, int64_t n, int per_row) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    Val<W> y, d, r;
    load(rec_at<W>(Y, i + j * ldy), y);
    load(rec_at<W>(D, per_row ? i : j), d);
    mul(y, d, r);
    store(rec_at<W>(Z, i + j * ldz), r);
}

// H (K x K lower, ld K) from the packed-upper Gram and the weights
template <int W>
__global__ void sampled_entries_kernel(const uint32_t *G, const uint32_t *w, int64_t K,
                                       uint32_t *H) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * K) return;
    const int64_t p = e % K, q = e / K;
    Val<W> r;
    if (p < q) {
        set_zero(r);
        store(rec_at<W>(H, e), r);
        return;
    }
    Val<W> wp, wq, g, t1, u1;
    load(rec_at<W>(w, p), wp);
    load(rec_at<W>(w, q), wq);
    load(rec_at<W>(G, p * (p + 1) / 2 + q), g);
    mul(wp, wq, t1);
    mul(g, g, u1);
    // RN(RN(t1 / 2) RN(u1 + u1)) = RN(t1 u1): both steps are exact scalings
    // by powers of two (entry(), sampled/mod.rs:1973-1988)
    mul(t1, u1, r);
    store(rec_at<W>(H, e), r);
}

// ---------------------------------------------------------------------------
// dots
// ---------------------------------------------------------------------------
// Window of SLOTS 32-bit positions: products of 2W words plus 16 words of
// exponent spread and carry headroom.
template <int W>
struct Win {
    static constexpr int SLOTS = 2 * W + 20;
};

// ---------------------------------------------------------------------------
// group dots: one exact dot shared by the G lanes of an aligned lane group
// ---------------------------------------------------------------------------
// Terms are indexed by their absolute position k in [k0, k1); lane gl of the
// group takes the k with k % G == gl, so a value written at position k by
// lane k % G is only ever read back by that same lane.  Every lane builds a
// partial slot window over the group's common base; the windows are summed
// across the group (integer adds of the unresolved slots: the same sum a
// single lane would hold after all k1 - k0 terms), and every lane then
// carries and rounds the identical total, so all lanes hold the result.
// LDS slot windows: thread t's slot k is the int64 at slots_sh[k * STRIDE + t],
// so slot k of the 16 lanes of a group is one contiguous 128-byte run.  The
// routines below take LDS-typed pointers so every access is a ds_ operation.
typedef __attribute__((address_space(3))) int64_t lds_i64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef long long i64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) i64x2 lds_i64x2;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

// s * (p << off) into this lane's window (ls: its slot 0; off >= 0 bits
// above the window base): L + 1 slots, all read, then all written.
template <int L, int STRIDE>
__device__ __forceinline__ void lds_add_words(lds_i64 *ls, int64_t off, const uint32_t (&p)[L], bool neg) {
    const int wo = (int)(off >> 5), bo = (int)(off & 31);
    lds_i64 *s = ls + (int64_t)wo * STRIDE;
    int64_t cur[L + 1];
#pragma unroll
    for (int i = 0; i <= L; i++) cur[i] = s[i * STRIDE];
#pragma unroll
    for (int i = 0; i <= L; i++) {
        const uint32_t hi = i < L ? p[i] : 0u, lo = i ? p[i - 1] : 0u;
        const uint32_t v = bo ? funnel(hi, lo, 32 - bo) : hi;  // (p << bo) word i
        s[i * STRIDE] = neg ? cur[i] - (int64_t)v : cur[i] + (int64_t)v;
    }
}

// s * (a * b) into this lane's window whose slot 0 sits at scale base.
template <int W, int STRIDE>
__device__ __forceinline__ void lds_add_product(lds_i64 *ls, int64_t base, const Val<W> &a, const Val<W> &b,
                                                bool negate) {
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
    lds_add_words<2 * W, STRIDE>(ls, (int64_t)a.e + b.e - 64 * (int64_t)W - base, p, neg);
}

// Sum the G windows of a lane group (slots [0, used); higher slots count as
// zero), carry-resolve the total and round it once into out; every lane of
// the group ends with the result.  Lane s % G sums slot s over the group in
// 16-byte reads and writes the total into its own slot s / G; every lane
// then reads all totals back, resolves the carries and writes the 32-bit
// words to the group's first slot rows (2G words a row), from which the
// words around the top are gathered.  The LDS operations of a wave complete
// in issue order, so each read above precedes the writes that reuse its
// location.
template <int W, int G, int SLOTS, int STRIDE>
__device__ __forceinline__ void group_finish(lds_i64 *ls, int gl, int used, int64_t base, Val<W> &out) {
    constexpr int ROWS = (SLOTS + G - 1) / G;
    constexpr int WROW = 2 * G;  // 32-bit words per slot row
    lds_i64 *g0 = ls - gl;       // slot 0 of the group's first lane
    int64_t tot[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const int sl = gl + G * r;
        int64_t acc = 0;
        if (sl < used) {
            const lds_i64x2 *row = (const lds_i64x2 *)(g0 + sl * STRIDE);
#pragma unroll
            for (int q = 0; q < G / 2; q++) {
                const i64x2 x = row[q];
                acc += x.x + x.y;
            }
        }
        tot[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < ROWS; r++) ls[r * STRIDE] = tot[r];
    uint32_t w[SLOTS];
    int64_t carry = 0;
#pragma unroll
    for (int k = 0; k < SLOTS; k += 2) {
        const i64x2 x = *(const lds_i64x2 *)(g0 + (k / G) * STRIDE + (k % G));
        int64_t t = x.x + carry;
        w[k] = (uint32_t)(uint64_t)t;
        carry = t >> 32;
        if (k + 1 < SLOTS) {
            t = x.y + carry;
            w[k + 1] = (uint32_t)(uint64_t)t;
            carry = t >> 32;
        }
    }
    const bool neg = carry < 0;
    if (neg) {
        uint64_t c = 1;
#pragma unroll
        for (int k = 0; k < SLOTS; k++) {
            const uint64_t t = (uint64_t)(~w[k]) + c;
            w[k] = (uint32_t)t;
            c = t >> 32;
        }
    }
    int top = -1, low = SLOTS;
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
        if (w[k]) top = k;
        if (w[SLOTS - 1 - k]) low = SLOTS - 1 - k;
    }
    if (top < 0) {
        set_zero(out);
        return;
    }
#pragma unroll
    for (int q = 0; q < (SLOTS + 3) / 4; q++) {
        const int k = 4 * q;
        lds_u32x4 *dst = (lds_u32x4 *)((lds_u32 *)(g0 + (k / WROW) * STRIDE) + (k % WROW));
        *dst = u32x4{w[k], k + 1 < SLOTS ? w[k + 1] : 0u, k + 2 < SLOTS ? w[k + 2] : 0u,
                     k + 3 < SLOTS ? w[k + 3] : 0u};
    }
    constexpr int GW = W + 2;
    const int lo = top - GW + 1;
    uint32_t g[GW + 1];
    g[0] = low < lo ? 1u : 0u;  // one sticky bit below the gathered words
#pragma unroll
    for (int i = 0; i < GW; i++) {
        const int k = lo + i;
        g[i + 1] = k >= 0 ? *((lds_u32 *)(g0 + (k / WROW) * STRIDE) + (k % WROW)) : 0u;
    }
    round_norm<W, GW + 1>(neg, g, base + 32 * (int64_t)(lo - 1), out);  // g[GW] = w[top] != 0
}

// group_finish for a whole-wave group (G = 64, SLOTS <= 64): lane s sums
// slot s over the wave, the window is carried and rounded on the wave
// (wmp::round_window: fixed carry steps, the top slot keeping the sign,
// round and sticky bits from ballots) and read into registers.  Same
// result as group_finish.
template <int W, int SLOTS>
__device__ __forceinline__ void wave_round(int64_t T, int gl, int64_t base, Val<W> &out);

// Lane s's total of slot s over the wave's 64 windows (0 for s >= used).
template <int STRIDE>
__device__ __forceinline__ int64_t window_total(lds_i64 *ls, int gl, int used) {
    lds_i64 *g0 = ls - gl;  // slot 0 of lane 0
    int64_t T = 0;
    if (gl < used) {
        const lds_i64x2 *row = (const lds_i64x2 *)(g0 + gl * STRIDE);
#pragma unroll
        for (int q = 0; q < 32; q++) {
            const i64x2 x = row[q];
            T += x.x + x.y;
        }
    }
    return T;
}

template <int W, int SLOTS, int STRIDE>
__device__ __forceinline__ void wave_finish(lds_i64 *ls, int gl, int used, int64_t base, Val<W> &out) {
    static_assert(SLOTS <= 64, "one slot a lane");
    wave_round<W, SLOTS>(window_total<STRIDE>(ls, gl, used), gl, base, out);
}

// The rounding half of wave_finish: lane s holds slot s's total (lanes at
// or above SLOTS hold zero); every lane ends with RN of the window's value
// at scale base (the window rounded on the wave, then read into registers).
template <int W, int SLOTS>
__device__ __forceinline__ void wave_round(int64_t T, int gl, int64_t base, Val<W> &out) {
    (void)gl;
    out = wmp::to_val(wmp::round_window<W, SLOTS>(T, base));
}

// The lane windows of a group dot before its finish: slots [0, used) of
// each lane cleared, then the lane's terms (k % G == gl) added at window
// base lo.  Every term must lie in the window.
template <int W, int G, int STRIDE>
__device__ __forceinline__ void group_accumulate(int64_t k0, int64_t k1, int gl, const uint32_t *xb, int64_t xs,
                                                 const uint32_t *yb, int64_t ys, bool negate_terms, lds_i64 *ls,
                                                 int used, int64_t lo) {
    auto fx = [&](int64_t k) { return xb + k * xs * (W + 2); };
    auto fy = [&](int64_t k) { return yb + k * ys * (W + 2); };
    const int64_t kf = k0 + (((int64_t)gl - k0) % G + G) % G;  // first k of this lane
    for (int sl = 0; sl < used; sl++) ls[sl * STRIDE] = 0;
    Val<W> xn, yn;
    int64_t k = kf;
    if (k < k1) {
        load(fx(k), xn);
        load(fy(k), yn);
    }
    for (; k < k1; k += G) {
        const Val<W> x = xn, y = yn;
        if (k + G < k1) {  // next term in flight during this one
            load(fx(k + G), xn);
            load(fy(k + G), yn);
        }
        if (is_zero(x) || is_zero(y)) continue;
        lds_add_product<W, STRIDE>(ls, lo, x, y, negate_terms);
    }
}

// Terms x_k = xb[k * xs], y_k = yb[k * ys] (record strides): every panel
// kernel's operands are affine in k.  Inlined at each call site, so the
// operands and the result stay in registers.  Groups are 4, 8, 16, 32 or 64
// lanes.
template <int W, int G, int STRIDE>
__device__ __forceinline__ bool group_dot(int64_t k0, int64_t k1, int gl,
                                                    const uint32_t *xb, int64_t xs,
                                                    const uint32_t *yb, int64_t ys, const Val<W> *c,
                                                    bool negate_terms, int64_t *slots, Val<W> &out,
                                                    int64_t hint_lo, int64_t hint_hi) {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "a group is an aligned part of one wave");
    constexpr int SLOTS = Win<W>::SLOTS;
    auto fx = [&](int64_t k) { return xb + k * xs * (W + 2); };
    auto fy = [&](int64_t k) { return yb + k * ys * (W + 2); };
    const int64_t kf = k0 + (((int64_t)gl - k0) % G + G) % G;  // first k of this lane
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (hint_lo <= hint_hi && (hint_hi - hint_lo) / 32 + 2 * W + 4 <= SLOTS) {
        lo = hint_lo;
        hi = hint_hi;
    } else {
        for (int64_t k = kf; k < k1; k += G) {
            const uint32_t *x = fx(k), *y = fy(k);
            const int32_t kx = (int32_t)x[W], ky = (int32_t)y[W];
            if (kx == mpd::KIND_ZERO || kx == -mpd::KIND_ZERO || ky == mpd::KIND_ZERO ||
                ky == -mpd::KIND_ZERO)
                continue;
            const int64_t sc = (int64_t)(int32_t)x[W + 1] + (int32_t)y[W + 1] - 64 * (int64_t)W;
            lo = min(lo, sc);
            hi = max(hi, sc);
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            lo = min(lo, (int64_t)__shfl_xor(lo, off, 64));
            hi = max(hi, (int64_t)__shfl_xor(hi, off, 64));
        }
    }
    bool any = lo != INT64_MAX;
    const bool cnz = c && !is_zero(*c);
    if (cnz) {
        const int64_t sc = (int64_t)c->e - 32 * W;
        lo = any ? min(lo, sc) : sc;
        hi = any ? max(hi, sc - 32 * W) : sc - 32 * W;
        any = true;
    }
    if (!any) {
        set_zero(out);
        return true;
    }
    if ((hi - lo) / 32 + 2 * W + 4 > SLOTS) return false;
    // terms land in slots [0, (hi - lo)/32 + 2W]
    const int used = (int)min((int64_t)SLOTS, (hi - lo) / 32 + 2 * W + 1);
    lds_i64 *ls = (lds_i64 *)slots;
    group_accumulate<W, G, STRIDE>(k0, k1, gl, xb, xs, yb, ys, negate_terms, ls, used, lo);
    if (cnz && gl == 0) lds_add_words<W, STRIDE>(ls, (int64_t)c->e - 32 * W - lo, c->m, c->kind < 0);
    if constexpr (G == 64 && SLOTS <= 64)
        wave_finish<W, SLOTS, STRIDE>(ls, gl, used, lo, out);
    else
        group_finish<W, G, SLOTS, STRIDE>(ls, gl, used, lo, out);
    return true;
}

// group_dot for a whole wave (G = 64) without a subtrahend, its result
// spread over the wave (wavemp.h): the same window, terms and rounding.
template <int W, int STRIDE>
__device__ __forceinline__ bool wave_dot(int64_t k0, int64_t k1, int gl, const uint32_t *xb, int64_t xs,
                                         const uint32_t *yb, int64_t ys, int64_t *slots, wmp::WV<W> &out,
                                         int64_t hint_lo, int64_t hint_hi) {
    constexpr int SLOTS = Win<W>::SLOTS;
    static_assert(SLOTS <= 64, "one slot a lane");
    auto fx = [&](int64_t k) { return xb + k * xs * (W + 2); };
    auto fy = [&](int64_t k) { return yb + k * ys * (W + 2); };
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (hint_lo <= hint_hi && (hint_hi - hint_lo) / 32 + 2 * W + 4 <= SLOTS) {
        lo = hint_lo;
        hi = hint_hi;
    } else {
        const int64_t kf = k0 + (((int64_t)gl - k0) % 64 + 64) % 64;
        for (int64_t k = kf; k < k1; k += 64) {
            const uint32_t *x = fx(k), *y = fy(k);
            const int32_t kx = (int32_t)x[W], ky = (int32_t)y[W];
            if (kx == mpd::KIND_ZERO || kx == -mpd::KIND_ZERO || ky == mpd::KIND_ZERO ||
                ky == -mpd::KIND_ZERO)
                continue;
            const int64_t sc = (int64_t)(int32_t)x[W + 1] + (int32_t)y[W + 1] - 64 * (int64_t)W;
            lo = min(lo, sc);
            hi = max(hi, sc);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, (int64_t)__shfl_xor(lo, off, 64));
            hi = max(hi, (int64_t)__shfl_xor(hi, off, 64));
        }
    }
    if (lo == INT64_MAX) {
        out = wmp::zero<W>();
        return true;
    }
    if ((hi - lo) / 32 + 2 * W + 4 > SLOTS) return false;
    const int used = (int)min((int64_t)SLOTS, (hi - lo) / 32 + 2 * W + 1);
    lds_i64 *ls = (lds_i64 *)slots;
    group_accumulate<W, 64, STRIDE>(k0, k1, gl, xb, xs, yb, ys, false, ls, used, lo);
    out = wmp::round_window<W, SLOTS>(window_total<STRIDE>(ls, gl, used), lo);
    return true;
}

// Lanes per exact dot in the panel kernels.  The row kernels (forward rows,
// trsm bases) run one wave per workgroup so their many small workgroups
// spread over all CUs; ldl_dots runs a whole block per workgroup (512
// threads, 256 at 512 bits).  The slot windows live in LDS: SLOTS x 8 bytes
// per thread.
constexpr int DOT_G = 16;
constexpr int ROW_THREADS = 64;

// Y = L^-1 B per column (forward_suffix_rows): L K x K lower (ld K),
// first[K] the row envelope, Y K x t (ld K) in place.  One lane group per
// column; y_i is stored by lane i % G, the only lane that reads it back.
// window hint from two exponent ranges (min > max: empty)
template <int W>
__device__ __forceinline__ void range_hint(int32_t amn, int32_t amx, int32_t bmn, int32_t bmx,
                                           int64_t &hl, int64_t &hh) {
    hl = INT64_MAX;
    hh = INT64_MIN;
    if (amn <= amx && bmn <= bmx) {
        hl = (int64_t)amn + bmn - 64 * W;
        hh = (int64_t)amx + bmx - 64 * W;
    }
}
__device__ __forceinline__ void range_add(int32_t &mn, int32_t &mx, int32_t kind, int32_t e) {
    if (kind != mpd::KIND_ZERO && kind != -mpd::KIND_ZERO) {
        mn = min(mn, e);
        mx = max(mx, e);
    }
}

// exponent range (min, max) of the non-zero entries of each row of the
// strictly lower triangle of L (mode 0: row j over columns k < j) or of each
// column (mode 1: column j over rows k > j); empty gives (INT32_MAX, INT32_MIN)
// one wave a row (column): the lanes stride the entries, then a wave min/max
template <int W>
__global__ __launch_bounds__(64) void exp_ranges_kernel(const uint32_t *L, int64_t ldl, int64_t n, int mode,
                                                        int2 *out) {
    const int64_t j = blockIdx.x;
    const int lane = threadIdx.x;
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    const int64_t k0 = mode == 0 ? 0 : j + 1, k1 = mode == 0 ? j : n;
    for (int64_t k = k0 + lane; k < k1; k += 64) {
        const uint32_t *rec = rec_at<W>(L, mode == 0 ? j + k * ldl : k + j * ldl);
        range_add(mn, mx, (int32_t)rec[W], (int32_t)rec[W + 1]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_xor(mn, off, 64));
        mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if (lane == 0) out[j] = make_int2(mn, mx);
}

template <int W, int G>
__global__ __launch_bounds__(ROW_THREADS) void forward_rows_kernel(const uint32_t *L_all, const int32_t *first_all,
                                                                   uint32_t *Y_all, const int2 *lr_all,
                                                                   const FwdLeaf *leaves, int64_t t, int *fail) {
    extern __shared__ int64_t slots_sh[];
    const FwdLeaf lf = leaves[blockIdx.y];
    const int gl = threadIdx.x & (G - 1);
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (c >= t) return;  // whole group
    const int64_t K = lf.k;
    const uint32_t *L = rec_at<W>(L_all, lf.l_off);
    const int32_t *first = first_all + lf.first_off;
    const int2 *lr = lr_all ? lr_all + lf.range_off : nullptr;
    int64_t *slots = slots_sh + threadIdx.x;
    uint32_t *col = rec_at<W>(Y_all, lf.y_off + c * K);
    int32_t ymn = INT32_MAX, ymx = INT32_MIN;  // y_k so far
    for (int64_t i = 0; i < K; i++) {
        Val<W> b, r;
        load(rec_at<W>(col, i), b);
        // y_i = RN(b_i - sum_{k in [first_i, i)} L_ik y_k)
        int64_t hl = INT64_MAX, hh = INT64_MIN;
        if (lr) range_hint<W>(lr[i].x, lr[i].y, ymn, ymx, hl, hh);
        if (!group_dot<W, G, ROW_THREADS>(first[i], i, gl, rec_at<W>(L, i), K, col, 1, &b, true, slots, r,
                                          hl, hh)) {
            if (gl == 0) atomicOr(fail, 1);
            return;
        }
        range_add(ymn, ymx, r.kind, r.e);
        if (gl == (int)(i % G)) store(rec_at<W>(col, i), r);
    }
}

// X L^T = B rows (solve_unit_lower_t base, arrow.rs:425-447) right-looking:
// one workgroup a row, each column c < n (n <= 128) with its own exact slot
// window in LDS holding the terms x_k L_ck, k < c.  Wave 0 runs the chain:
// step j takes column j's window, adds the one term x_{j-1} L_{j,j-1} in
// registers, rounds once (the exact dot of x_k L_jk over k < j) and forms
// x_j = RN(b_j - t) on the wave (wavemp.h).  Waves 1 and 2 (thread c - 64
// owning column c) add the other terms one step behind: in the interval of
// chain step j they add x_{j-1} L_{c,j-1} for c > j.  One barrier an
// interval, so a step costs the chain's rounding, sum and one product while
// the per-column products run beside it.  A window's base is set at its
// first term (by its column's thread, or by the chain when that term is the
// last one) from the exponent range of L's row (rows of L over k < c) with
// `room` bits (128 by default) for x to shrink; a term outside the window
// puts the row on the fallback list (fb_rows) with B's row untouched, and
// the left-looking kernel then solves it.  x_j is kept in column j's window
// once the chain has read it (word w in the low or high half of slot w / 2),
// so LDS holds only the windows and their bases.  Results are those of the
// left-looking kernel: the same exact dots, each rounded once.
constexpr int RL_COLS = 128;               // columns (window slots per slot row)
// wave priority for the latency chains of trsm_rl_kernel and ldl_rows_kernel
constexpr bool g_chain_prio = true;
constexpr int RL_THREADS = RL_COLS + 64;   // the chain wave, then one thread a column

template <int W>
__device__ __forceinline__ int64_t rl_base(int32_t xe, int32_t le, const int2 *lr, int64_t c, int64_t room) {
    const int64_t sc = (int64_t)xe + le - 64 * W;
    return lr ? (int64_t)xe - room + lr[c].x - 64 * W : sc - 2 * room;
}

template <int W>
__global__ __launch_bounds__(RL_THREADS) void trsm_rl_kernel(
    const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B, int64_t ldb, int64_t m,
    const int32_t *low, const int2 *lr, int64_t room, int *fb_count, int32_t *fb_rows) {
    constexpr int SLOTS = Win<W>::SLOTS;
    static_assert(SLOTS <= 64, "one slot a lane");
    // [SLOTS][RL_COLS] windows, then the window bases [RL_COLS]
    extern __shared__ int64_t slots_sh[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t i = blockIdx.x;
    if (i >= m) return;  // the whole workgroup
    lds_i64 *win = (lds_i64 *)slots_sh;
    int64_t *bases = slots_sh + SLOTS * RL_COLS;
    // word w of x_c, in column c's window after the chain has consumed it
    auto xword = [&](int64_t c, int w) -> lds_u32 * { return (lds_u32 *)(win + (w >> 1) * RL_COLS + c) + (w & 1); };
    const int64_t lo = low[i];
    // a term outside its window: ovf_step[j & 1] stops every wave after the
    // barrier of step j (two slots, so a wave already in step j + 1 never
    // writes the slot another wave is about to read); ovf_any for the end
    __shared__ int ovf_step[2], ovf_any;
    for (int e = tid; e < SLOTS * RL_COLS; e += RL_THREADS) win[e] = 0;
    for (int e = tid; e < RL_COLS; e += RL_THREADS) bases[e] = INT64_MIN;
    if (tid == 0) {
        ovf_step[0] = ovf_step[1] = 0;
        ovf_any = 0;
    }
    __syncthreads();
    bool ovf = false;
    // the waves exchange only LDS data inside the loop: a barrier with LDS fences
    auto step_end = [&](int64_t j) -> bool {
        if (ovf) {
            ovf_step[j & 1] = 1;
            ovf_any = 1;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        return ovf_step[j & 1] != 0;
    };
    if (wave == 0) {  // the chain
        // the chain's instructions first where workgroups share a SIMD
        if (g_chain_prio) __builtin_amdgcn_s_setprio(2);
        wmp::WV<W> xprev = wmp::zero<W>();
        // b_j and L_{j,j-1} one step ahead
        uint32_t bn = lo < n ? wmp::load_word<W>(rec_at<W>((const uint32_t *)B, i + lo * ldb)) : 0u, ln = 0u;
        for (int64_t j = lo; j < n; j++) {
            const uint32_t bw = bn, lw = ln;
            if (j + 1 < n) {
                bn = wmp::load_word<W>(rec_at<W>((const uint32_t *)B, i + (j + 1) * ldb));
                ln = wmp::load_word<W>(rec_at<W>(L, (j + 1) + j * ldl));
            }
            const wmp::WV<W> bj = wmp::from_word<W>(bw);
            const wmp::WV<W> lj = j > lo ? wmp::from_word<W>(lw) : wmp::zero<W>();
            int64_t T = lane < SLOTS ? (int64_t)win[lane * RL_COLS + j] : 0;
            int64_t base = bases[j];
            if (!wmp::is_zero(xprev) && !wmp::is_zero(lj)) {  // the last term x_{j-1} L_{j,j-1}
                if (base == INT64_MIN) base = rl_base<W>(xprev.e, lj.e, lr, j, room);
                const int64_t off = (int64_t)xprev.e + lj.e - 64 * W - base;
                if (off < 0 || off / 32 + 2 * W + 4 > SLOTS)
                    ovf = true;
                else
                    T = wmp::add_product<W, SLOTS>(T, xprev, lj, off, false);
            }
            const wmp::WV<W> t = base == INT64_MIN ? wmp::zero<W>() : wmp::round_window<W, SLOTS>(T, base);
            xprev = wmp::add(bj, t, true);  // x_j = RN(b_j - t)
            if (lane < W + 2)  // after the window's reads above (LDS operations of a wave complete in order)
                *xword(j, lane) = lane < W ? xprev.w : (lane == W ? (uint32_t)xprev.kind : (uint32_t)xprev.e);
            if (step_end(j)) break;
        }
    } else {  // thread c adds x_k L_ck for c > k + 1, one step behind the chain
        const int c = tid - 64;
        int64_t base = INT64_MIN;
        Val<W> lv, ln;
        auto load_l = [&](int64_t k, Val<W> &dst) {
            if (c > k + 1 && c < n) load(rec_at<W>(L, c + k * ldl), dst);
            else set_zero(dst);
        };
        if (lo < n) load_l(lo, ln);
        for (int64_t j = lo; j < n; j++) {
            const int64_t k = j - 1;  // the chain's previous step
            if (k >= lo) {
                lv = ln;
                if (k + 1 < n) load_l(k + 1, ln);
                if (!is_zero(lv)) {  // zero for c <= k + 1 and c >= n
                    Val<W> x;
#pragma unroll
                    for (int w = 0; w < W; w++) x.m[w] = *xword(k, w);
                    x.kind = (int32_t)*xword(k, W);
                    x.e = (int32_t)*xword(k, W + 1);
                    if (!is_zero(x)) {
                        if (base == INT64_MIN) {
                            base = rl_base<W>(x.e, lv.e, lr, c, room);
                            bases[c] = base;
                        }
                        const int64_t off = (int64_t)x.e + lv.e - 64 * W - base;
                        if (off < 0 || off / 32 + 2 * W + 4 > SLOTS)
                            ovf = true;
                        else
                            lds_add_product<W, RL_COLS>(win + c, base, x, lv, false);
                    }
                }
            }
            if (step_end(j)) break;
        }
    }
    __syncthreads();
    if (ovf_any) {
        if (tid == 0) fb_rows[atomicAdd(fb_count, 1)] = (int32_t)i;
        return;
    }
    for (int64_t e = lo * (W + 2) + tid; e < n * (W + 2); e += RL_THREADS) {  // B's row: x_c for c in [lo, n)
        const int64_t c = e / (W + 2);
        const int w = (int)(e - c * (W + 2));
        B[(size_t)(i + c * ldb) * (W + 2) + w] = *xword(c, w);
    }
}

// X L^T = B rows (solve_unit_lower_t base): B m x n (ldb), L n x n (ldl,
// unit lower, diagonal unread), low[m]; one lane group per row.
template <int W, int G>
__global__ __launch_bounds__(ROW_THREADS) void trsm_base_kernel(
    const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B, int64_t ldb, int64_t m,
    const int32_t *low, const int2 *lr, int *fail, const int32_t *rows, const int *nrows) {
    extern __shared__ int64_t slots_sh[];
    const int gl = threadIdx.x & (G - 1);
    int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (rows) {  // only the listed rows
        if (i >= *nrows) return;
        i = rows[i];
    }
    if (i >= m) return;
    int64_t *slots = slots_sh + threadIdx.x;
    const int64_t lo = low[i];
    int32_t xmn = INT32_MAX, xmx = INT32_MIN;  // x_ik so far
    for (int64_t j = lo; j < n; j++) {
        Val<W> t, x, r;
        int64_t hl = INT64_MAX, hh = INT64_MIN;
        if (lr) range_hint<W>(xmn, xmx, lr[j].x, lr[j].y, hl, hh);
        if (!group_dot<W, G, ROW_THREADS>(lo, j, gl, rec_at<W>((const uint32_t *)B, i), ldb, rec_at<W>(L, j),
                                          ldl, nullptr, false, slots, t, hl, hh)) {
            if (gl == 0) atomicOr(fail, 1);
            return;
        }
        load(rec_at<W>(B, i + j * ldb), x);
        add_signed(x, t, true, r);
        range_add(xmn, xmx, r.kind, r.e);
        if (gl == (int)(j % G)) store(rec_at<W>(B, i + j * ldb), r);
    }
}

// X L = B rows for unit lower L, descending columns: x_j = RN(x_j -
// RN(sum_{k > j} x_k L_kj)); B m x n (ldb), L n x n (ldl); one lane group
// per row.
template <int W, int G>
__global__ __launch_bounds__(ROW_THREADS) void trsm_base_n_kernel(
    const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B, int64_t ldb, int64_t m,
    const int2 *lc, int *fail) {
    extern __shared__ int64_t slots_sh[];
    const int gl = threadIdx.x & (G - 1);
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (i >= m) return;
    int64_t *slots = slots_sh + threadIdx.x;
    int32_t xmn = INT32_MAX, xmx = INT32_MIN;  // x_ik so far (k > j)
    for (int64_t j = n - 1; j >= 0; j--) {
        Val<W> t, x, r;
        int64_t hl = INT64_MAX, hh = INT64_MIN;
        if (lc) range_hint<W>(xmn, xmx, lc[j].x, lc[j].y, hl, hh);
        if (!group_dot<W, G, ROW_THREADS>(j + 1, n, gl, rec_at<W>((const uint32_t *)B, i), ldb,
                                          rec_at<W>(L, j * ldl), 1, nullptr, false, slots, t, hl, hh)) {
            if (gl == 0) atomicOr(fail, 1);
            return;
        }
        load(rec_at<W>(B, i + j * ldb), x);
        add_signed(x, t, true, r);
        range_add(xmn, xmx, r.kind, r.e);
        if (gl == (int)(j % G)) store(rec_at<W>(B, i + j * ldb), r);
    }
}

// out (n x m, ldo) = in^T (in: m x n, ldi), values
// trsm_base_kernel / trsm_base_n_kernel for one whole wave a row and
// n <= 128: the row's solved x_k stay in registers (lane k % 64 holds x_k),
// so each step's dot reads only L's entries (fetched a step ahead) and
// rounds on the wave (wavemp.h); x_j = RN(b_j - t) is formed on the wave
// too.  Forward (solve_unit_lower_t): j = lo .. n - 1 over terms
// x_k L(j, k), k in [lo, j), window hint from the x range and lr[j] (L's row
// range); backward: j = n - 1 .. 0 over x_k L(k, j), k in (j, n), hint from
// lr[j] (L's column range).  Same windows, terms and roundings as the group
// kernels; a dot beyond its window sets *fail.
template <int W, bool BACK>
__device__ __forceinline__ void trsm_row_wave(const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B, int64_t ldb,
                                              int64_t i, int64_t lo, const int2 *lr, int *fail, int64_t *slots_sh) {
    constexpr int SLOTS = Win<W>::SLOTS;
    static_assert(SLOTS <= 64, "one slot a lane");
    const int lane = threadIdx.x & 63;
    lds_i64 *ls = (lds_i64 *)(slots_sh + lane);
    Val<W> x0, x1;  // x_k for k = lane, lane + 64
    set_zero(x0);
    set_zero(x1);
    int32_t xmn = INT32_MAX, xmx = INT32_MIN;  // x_k solved so far
    // the step's terms: k in [lo, j) forward, (j, n) backward
    auto in_dot = [&](int64_t j, int64_t k) { return BACK ? (k > j && k < n) : (k >= lo && k < j); };
    auto load_l = [&](int64_t j, int64_t k, Val<W> &dst) {
        if (in_dot(j, k)) load(rec_at<W>(L, BACK ? k + j * ldl : j + k * ldl), dst);
        else set_zero(dst);
    };
    const int64_t j0 = BACK ? n - 1 : lo, jend = BACK ? -1 : n, dj = BACK ? -1 : 1;
    Val<W> l0n, l1n;
    uint32_t bn = 0;
    if (j0 != jend) {
        load_l(j0, lane, l0n);
        load_l(j0, lane + 64, l1n);
        bn = wmp::load_word<W>(rec_at<W>((const uint32_t *)B, i + j0 * ldb));
    }
    for (int64_t j = j0; j != jend; j += dj) {
        const Val<W> l0 = l0n, l1 = l1n;
        const uint32_t bw = bn;
        if (j + dj != jend) {
            load_l(j + dj, lane, l0n);
            load_l(j + dj, lane + 64, l1n);
            bn = wmp::load_word<W>(rec_at<W>((const uint32_t *)B, i + (j + dj) * ldb));
        }
        const bool t0 = !is_zero(x0) && !is_zero(l0), t1 = !is_zero(x1) && !is_zero(l1);
        int64_t hl = INT64_MAX, hh = INT64_MIN;
        if (lr) range_hint<W>(xmn, xmx, lr[j].x, lr[j].y, hl, hh);
        if (!(hl <= hh && (hh - hl) / 32 + 2 * W + 4 <= SLOTS)) {  // the terms' own range
            hl = INT64_MAX;
            hh = INT64_MIN;
            if (t0) {
                const int64_t sc = (int64_t)x0.e + l0.e - 64 * W;
                hl = min(hl, sc);
                hh = max(hh, sc);
            }
            if (t1) {
                const int64_t sc = (int64_t)x1.e + l1.e - 64 * W;
                hl = min(hl, sc);
                hh = max(hh, sc);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                hl = min(hl, (int64_t)__shfl_xor(hl, off, 64));
                hh = max(hh, (int64_t)__shfl_xor(hh, off, 64));
            }
        }
        wmp::WV<W> t = wmp::zero<W>();
        if (hl != INT64_MAX) {
            if ((hh - hl) / 32 + 2 * W + 4 > SLOTS) {
                if (lane == 0) atomicOr(fail, 1);
                return;
            }
            const int used = (int)min((int64_t)SLOTS, (hh - hl) / 32 + 2 * W + 1);
            for (int sl = 0; sl < used; sl++) ls[sl * 64] = 0;
            if (t0) lds_add_product<W, 64>(ls, hl, x0, l0, false);
            if (t1) lds_add_product<W, 64>(ls, hl, x1, l1, false);
            t = wmp::round_window<W, SLOTS>(window_total<64>(ls, lane, used), hl);
        }
        const wmp::WV<W> r = wmp::add(wmp::from_word<W>(bw), t, true);
        if (!wmp::is_zero(r)) {
            xmn = min(xmn, r.e);
            xmx = max(xmx, r.e);
        }
        wmp::store(rec_at<W>(B, i + j * ldb), r);
        const Val<W> rv = wmp::to_val(r);
        if (lane == (int)(j & 63)) {
            if (j < 64) x0 = rv;
            else x1 = rv;
        }
    }
}

template <int W>
__global__ __launch_bounds__(64) void trsm_base_wave_kernel(const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B,
                                                            int64_t ldb, int64_t m, const int32_t *low, const int2 *lr,
                                                            int *fail, const int32_t *rows, const int *nrows) {
    extern __shared__ int64_t slots_sh[];
    int64_t i = blockIdx.x;
    if (rows) {  // only the listed rows
        if (i >= *nrows) return;
        i = rows[i];
    }
    if (i >= m) return;
    trsm_row_wave<W, false>(L, ldl, n, B, ldb, i, low[i], lr, fail, slots_sh);
}

template <int W>
__global__ __launch_bounds__(64) void trsm_base_n_wave_kernel(const uint32_t *L, int64_t ldl, int64_t n, uint32_t *B,
                                                              int64_t ldb, int64_t m, const int2 *lc, int *fail) {
    extern __shared__ int64_t slots_sh[];
    const int64_t i = blockIdx.x;
    if (i >= m) return;
    trsm_row_wave<W, true>(L, ldl, n, B, ldb, i, 0, lc, fail, slots_sh);
}

template <int W>
__global__ void transpose_kernel(const uint32_t *in, int64_t ldi, int64_t m, int64_t n,
                                 uint32_t *out, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    const uint32_t *src = rec_at<W>(in, i + j * ldi);
    uint32_t *dst = rec_at<W>(out, j + i * ldo);
    for (int u = 0; u < W + 2; u++) dst[u] = src[u];
}

// ldl_dots on one block per workgroup.  Blocks: (offset into A, ld, n,
// offset into dinv, first[] offset, sign offset); r and w are per-block
// packed-lower scratch (n(n+1)/2 values each, at scratch + 2*tri(n)*block).
// Column j: lane group 0 computes row j + 1 and then the pivot of column
// j + 1 (it needs only that row), while the other groups compute rows
// j + 2 .. n - 1; one barrier per column orders every scratch write before
// its reads.  r_ij and w_ij are stored by lane j % G of their group, the lane
// that reads position j back in a later dot of the same group.  Running
// exponent ranges per row of r and w (LDS) give each dot its window without
// a range pass when they fit.
struct LdlBlock {
    int64_t a_off, ld, n, dinv_off, first_off, sign_off, scratch_off;
};

constexpr int LDL_STATS_MAX = 256;  // rows with exponent ranges in LDS
template <int W>
constexpr int ldl_threads() { return W >= 16 ? 256 : 512; }

template <int W, int G>
__global__ __launch_bounds__(ldl_threads<W>()) void ldl_dots_kernel(
    uint32_t *A, uint32_t *dinv, const int32_t *first_all, const int8_t *signs, int sign_mode,
    const LdlBlock *blocks, uint32_t *scratch, const uint32_t *reg_eps, const uint32_t *reg_delta,
    int *status) {
    constexpr int SLOTS = Win<W>::SLOTS;
    constexpr int TH = ldl_threads<W>();
    extern __shared__ int64_t slots_sh[];
    __shared__ int bad;
    const LdlBlock B = blocks[blockIdx.x];
    uint32_t *a = A + B.a_off * (W + 2);
    uint32_t *dv = dinv + B.dinv_off * (W + 2);
    const int32_t *first = first_all + B.first_off;
    const int64_t n = B.n, ld = B.ld;
    auto tri = [](int64_t i) { return i * (i + 1) / 2; };
    uint32_t *r = scratch + B.scratch_off * (W + 2);
    uint32_t *w = r + tri(n) * (W + 2);
    uint32_t *dstore = w + tri(n) * (W + 2);  // n pivots
    const int nthreads = blockDim.x;
    int64_t *slots = slots_sh + threadIdx.x;
    const int gl = threadIdx.x & (G - 1), grp = threadIdx.x / G, groups = nthreads / G;
    int32_t *ex = (int32_t *)(slots_sh + (size_t)SLOTS * nthreads);
    int32_t *rmin = ex, *rmax = ex + LDL_STATS_MAX, *wmin = ex + 2 * LDL_STATS_MAX,
            *wmax = ex + 3 * LDL_STATS_MAX;
    const bool stats = n <= LDL_STATS_MAX;
    if (stats)
        for (int64_t i = threadIdx.x; i < n; i += nthreads) {
            rmin[i] = wmin[i] = INT32_MAX;
            rmax[i] = wmax[i] = INT32_MIN;
        }
    __shared__ int stop;
    if (threadIdx.x == 0) {
        bad = 0;
        // a failed pivot or an earlier row kernel's window overflow stops the block
        stop = status[0] | status[2];
    }
    __syncthreads();
    if (stop != 0) return;  // the whole workgroup

    // window hint of the dot of r row i with w row j from the ranges so far
    auto hint = [&](int64_t i, int64_t j, int32_t rmn, int32_t rmx, int32_t wmn, int32_t wmx,
                    int64_t &hl, int64_t &hh) {
        hl = INT64_MAX;
        hh = INT64_MIN;
        if (stats && rmn <= rmx && wmn <= wmx) {
            hl = (int64_t)rmn + wmn - 64 * W;
            hh = (int64_t)rmx + wmx - 64 * W;
        }
    };
    // DenseLeaf::pivot of column j from its exact dot (group 0; lane 0 writes)
    auto pivot = [&](int64_t j, int64_t hl, int64_t hh) {
        const int64_t fj = first[j];
        const uint32_t *xr = rec_at<W>((const uint32_t *)r, tri(j));
        const uint32_t *yr = rec_at<W>((const uint32_t *)w, tri(j));
        Val<W> ajj, t, d;
        load(rec_at<W>(a, j + j * ld), ajj);
        bool ok = group_dot<W, G, TH>(fj, j, gl, xr, 1, yr, 1, nullptr, false, slots, t, hl, hh);
        if (ok) {
            add_signed(ajj, t, true, d);
            // exact retry of a cancelled pivot (uniform over the group)
            if (is_zero(d) && j > fj)
                ok = group_dot<W, G, TH>(fj, j, gl, xr, 1, yr, 1, &ajj, true, slots, d, hl, hh);
        }
        if (gl != 0) return;
        if (!ok) {
            bad = 2;  // window overflow: host path
            return;
        }
        const int sg = sign_mode == 0 ? 1 : (sign_mode == 1 ? signs[0] : signs[B.sign_off + j]);
        int fail = 0;
        if (!is_regular(d) && !is_zero(d)) fail = 1;
        if (!fail && reg_eps) {
            Val<W> ds = d, e, diff;
            if (sg < 0 && is_regular(ds)) ds.kind = -ds.kind;
            load(reg_eps, e);
            add_signed(ds, e, true, diff);
            if (!is_zero(diff) && diff.kind < 0) {
                load(reg_delta, d);
                if (sg < 0 && is_regular(d)) d.kind = -d.kind;
                atomicAdd(status + 1, 1);  // regularized count
            }
        }
        if (!fail && is_zero(d)) fail = 1;
        if (fail) {
            bad = 1;
            return;
        }
        store(rec_at<W>(dstore, j), d);
        Val<W> di;
        recip(d, di);
        store(rec_at<W>(dv, j), di);
    };
    // r_ij, w_ij of column j (a lane group); returns them in every lane
    auto entry = [&](int64_t i, int64_t j, const Val<W> &dj, const Val<W> &dij, Val<W> &v,
                     Val<W> &wv) -> bool {
        if (first[i] > j) {
            set_zero(v);
        } else {
            const int64_t lo = max(first[i], first[j]);
            int64_t hl, hh;
            if (stats)
                hint(i, j, rmin[i], rmax[i], wmin[j], wmax[j], hl, hh);
            else
                hint(i, j, 1, 0, 1, 0, hl, hh);
            Val<W> t, aij, diff;
            if (!group_dot<W, G, TH>(lo, j, gl, rec_at<W>((const uint32_t *)r, tri(i)), 1,
                                     rec_at<W>((const uint32_t *)w, tri(j)), 1, nullptr, false, slots,
                                     t, hl, hh))
                return false;
            load(rec_at<W>(a, i + j * ld), aij);
            add_signed(aij, t, true, diff);
            mul(diff, dij, v);
        }
        mul(v, dj, wv);
        if (gl == (int)(j % G)) {
            store(rec_at<W>(r, tri(i) + j), v);
            store(rec_at<W>(w, tri(i) + j), wv);
        }
        if (stats && gl == 0) {
            if (!is_zero(v)) {
                rmin[i] = min(rmin[i], v.e);
                rmax[i] = max(rmax[i], v.e);
            }
            if (!is_zero(wv)) {
                wmin[i] = min(wmin[i], wv.e);
                wmax[i] = max(wmax[i], wv.e);
            }
        }
        return true;
    };

    if (grp == 0 && n > 0) pivot(0, INT64_MAX, INT64_MIN);
    __syncthreads();
    for (int64_t j = 0; j < n && !bad; j++) {
        Val<W> dj, dij;
        load(rec_at<W>((const uint32_t *)dstore, j), dj);
        load(rec_at<W>((const uint32_t *)dv, j), dij);
        if (grp == 0) {
            if (j + 1 < n) {
                const int64_t i = j + 1;
                int32_t rmn = 1, rmx = 0, wmn = 1, wmx = 0;
                if (stats) {
                    rmn = rmin[i];
                    rmx = rmax[i];
                    wmn = wmin[i];
                    wmx = wmax[i];
                }
                Val<W> v, wv;
                if (!entry(i, j, dj, dij, v, wv)) {
                    if (gl == 0) atomicMax(&bad, 2);
                } else {
                    // the pivot's range: row i so far plus the entry just made
                    if (!is_zero(v)) {
                        rmn = min(rmn, v.e);
                        rmx = max(rmx, v.e);
                    }
                    if (!is_zero(wv)) {
                        wmn = min(wmn, wv.e);
                        wmx = max(wmx, wv.e);
                    }
                    int64_t hl, hh;
                    hint(i, i, rmn, rmx, wmn, wmx, hl, hh);
                    pivot(i, hl, hh);
                }
            }
        } else {
            for (int64_t i = j + 2 + (grp - 1); i < n; i += groups - 1) {
                Val<W> v, wv;
                if (!entry(i, j, dj, dij, v, wv)) {
                    if (gl == 0) atomicMax(&bad, 2);
                }
            }
        }
        __syncthreads();
    }
    if (bad) {
        if (threadIdx.x == 0) {
            atomicMax(status, bad);
            if (bad == 1) atomicMax(status + 3, (int)blockIdx.x + 1);  // a failing block, 1-based
        }
        return;
    }
    // write the unit lower factor
    for (int64_t e = threadIdx.x; e < n * n; e += nthreads) {
        const int64_t i = e % n, j = e / n;
        if (i < j) continue;
        Val<W> v;
        if (i == j) {
            for (int u = 0; u < W; u++) v.m[u] = 0;
            v.m[W - 1] = 0x80000000u;
            v.kind = mpd::KIND_REG;
            v.e = 1;
        } else {
            load(rec_at<W>((const uint32_t *)r, tri(i) + j), v);
        }
        store(rec_at<W>(a, i + j * ld), v);
    }
}

// ldl_dots over many workgroups (arrow.rs:156-248): launch j forms column j
// of every block with n > j + 1, one lane group per row i > j, and the group
// of row j + 1 then forms pivot j + 1 (its look-ahead); launch -1 forms
// pivot 0.  The kernel boundary between launches is the only
// synchronisation, so the rows of a column spread over as many CUs as there
// are rows.  State between launches lives in the block's scratch: the r and w
// rows (row-major triangles), the pivots d, and per-row exponent ranges
// (r min, r max, w min, w max) that bound each dot's window.  status[0]
// (1: failed pivot, 2: window overflow) stops every later launch.
template <int W>
struct LdlCol {
    uint32_t *a, *dv, *r, *w, *dstore;
    int4 *rng;
    int64_t *prow;  // [2][n][SLOTS + 1]: row windows by column parity (slot totals, then the base)
    int *flags;     // [2][n]: pivot_ready, progress (the one-launch form, ldl_rows_kernel)
    const int32_t *first;
    int64_t n, ld;
    static __device__ int64_t tri(int64_t i) { return i * (i + 1) / 2; }
    __device__ LdlCol(uint32_t *A, uint32_t *dinv, const int32_t *first_all, uint32_t *scratch,
                      const LdlBlock &B) {
        a = A + B.a_off * (W + 2);
        dv = dinv + B.dinv_off * (W + 2);
        first = first_all + B.first_off;
        n = B.n;
        ld = B.ld;
        r = scratch + B.scratch_off * (W + 2);
        w = r + tri(n) * (W + 2);
        dstore = w + tri(n) * (W + 2);
        rng = (int4 *)(dstore + n * (W + 2));
        prow = (int64_t *)(dstore + 2 * n * (W + 2));
        flags = (int *)(dstore + 22 * n * (W + 2));
    }
    // row i's window for column j (written by launch j - 1, read by launch j)
    __device__ int64_t *row_window(int64_t j, int64_t i) const {
        return prow + ((j & 1) * n + i) * (Win<W>::SLOTS + 1);
    }
};

// DenseLeaf::pivot (arrow.rs:618-634) of column j from the exact dot of
// its r and w rows (one lane group; lane 0 writes d), or from that dot's
// rounding when the caller has it.  Returns the status code (0 ok, 1 failed
// pivot, 2 window overflow) in every lane.
template <int W, int G>
__device__ int ldl_col_pivot(const LdlCol<W> &c, int64_t j, int gl, int64_t hl, int64_t hh,
                             int64_t *slots, const int8_t *signs, int sign_mode, int64_t sign_off,
                             const uint32_t *reg_eps, const uint32_t *reg_delta, int *status,
                             const Val<W> *dot = nullptr) {
    const int64_t fj = c.first[j];
    const uint32_t *xr = rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(j));
    const uint32_t *yr = rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j));
    Val<W> ajj, t, d;
    load(rec_at<W>(c.a, j + j * c.ld), ajj);
    bool ok = true;
    if (dot)  // RN of the exact dot, formed by the caller
        t = *dot;
    else
        ok = group_dot<W, G, ROW_THREADS>(fj, j, gl, xr, 1, yr, 1, nullptr, false, slots, t, hl, hh);
    if (ok) {
        add_signed(ajj, t, true, d);
        // a pivot cancelled to zero is the exact difference rounded once
        if (is_zero(d) && j > fj)
            ok = group_dot<W, G, ROW_THREADS>(fj, j, gl, xr, 1, yr, 1, &ajj, true, slots, d, hl, hh);
    }
    if (!ok) return 2;
    const int sg = sign_mode == 0 ? 1 : (sign_mode == 1 ? signs[0] : signs[sign_off + j]);
    if (!is_regular(d) && !is_zero(d)) return 1;
    if (reg_eps) {
        Val<W> ds = d, e, diff;
        if (sg < 0 && is_regular(ds)) ds.kind = -ds.kind;
        load(reg_eps, e);
        add_signed(ds, e, true, diff);
        if (!is_zero(diff) && diff.kind < 0) {
            load(reg_delta, d);
            if (sg < 0 && is_regular(d)) d.kind = -d.kind;
            if (gl == 0) atomicAdd(status + 1, 1);  // regularized count
        }
    }
    if (is_zero(d)) return 1;
    if (gl == 0) store(rec_at<W>(c.dstore, j), d);  // RN(1/d) is formed by the next launch
    return 0;
}

// ldl_col_pivot for a whole wave from the RN of the exact dot t spread over
// the wave, with a_jj (and the regularization's eps and delta) loaded
// beforehand.  Returns 3 when the pivot cancels to zero (the caller then
// takes ldl_col_pivot's whole-dot path, which rounds the exact difference
// once).
template <int W>
__device__ __forceinline__ int ldl_col_pivot_wave(const LdlCol<W> &c, int64_t j, const wmp::WV<W> &t,
                                                  const wmp::WV<W> &ajj, bool reg, const wmp::WV<W> &eps,
                                                  const wmp::WV<W> &delta, int sg, int *status) {
    wmp::WV<W> d = wmp::add(ajj, t, true);
    if (wmp::is_zero(d) && j > c.first[j]) return 3;
    if (!wmp::is_regular(d) && !wmp::is_zero(d)) return 1;
    if (reg) {
        wmp::WV<W> ds = d;
        if (sg < 0 && wmp::is_regular(ds)) ds.kind = -ds.kind;
        // RN(ds - eps) < 0 exactly when ds < eps
        const bool below = wmp::is_regular(eps) || wmp::is_zero(eps)
                               ? wmp::compare(ds, eps) < 0
                               : [&] {
                                     const wmp::WV<W> diff = wmp::add(ds, eps, true);
                                     return !wmp::is_zero(diff) && diff.kind < 0;
                                 }();
        if (below) {
            d = delta;
            if (sg < 0 && wmp::is_regular(d)) d.kind = -d.kind;
            if (wmp::lane() == 0) atomicAdd(status + 1, 1);  // regularized count
        }
    }
    if (wmp::is_zero(d)) return 1;
    wmp::store(rec_at<W>(c.dstore, j), d);  // RN(1/d) is formed by the next launch
    return 0;
}

// Launch j: the first wave of a workgroup holds the row lane groups; the
// second forms RN(1 / d_j) meanwhile (the reciprocal the rows' entries need
// only after their dots), hands it over through LDS and, in workgroup 0 of
// the block, stores it; in workgroup 0 (the one holding row j + 1, with
// whole-wave groups) the third accumulates pivot j + 1's exact dot over
// k < j into its lane windows, so after the entry (j + 1, j) only its one
// term is added before the rounding.  The window is the row's exponent
// range so far with 128 bits of room each side; a last term outside it, or
// no range, sends the pivot through the whole dot.  The status is read once
// per workgroup, so all waves take the same path to the barrier.
template <int W, int G>
__global__ __launch_bounds__(3 * ROW_THREADS) void ldl_col_kernel(
    uint32_t *A, uint32_t *dinv, const int32_t *first_all, const int8_t *signs, int sign_mode,
    const LdlBlock *blocks, uint32_t *scratch, const uint32_t *reg_eps, const uint32_t *reg_delta,
    int *status, int64_t j) {
    constexpr int SLOTS = Win<W>::SLOTS;
    extern __shared__ int64_t slots_sh[];  // [SLOTS][64] row groups, then [SLOTS][64] pivot windows
    __shared__ uint32_t dinv_sh[W + 2];
    __shared__ int stop;
    __shared__ int64_t pre_lo;
    if (threadIdx.x == 0) stop = status[0] | status[2];  // a failed pivot, or an earlier row kernel's window overflow
    __syncthreads();
    if (stop != 0) return;  // the whole workgroup
    const LdlBlock B = blocks[blockIdx.y];
    const LdlCol<W> c(A, dinv, first_all, scratch, B);
    const int wave = threadIdx.x / ROW_THREADS, lane = threadIdx.x & (ROW_THREADS - 1);
    const int gl = lane & (G - 1);
    const int64_t i = j + 1 + ((int64_t)blockIdx.x * ROW_THREADS + lane) / G;
    int64_t *slots = slots_sh + lane;
    lds_i64 *pre_win = (lds_i64 *)(slots_sh + SLOTS * ROW_THREADS) + lane;
    auto fail = [&](int code) {
        if (gl == 0) {
            atomicMax(status, code);
            if (code == 1) atomicMax(status + 3, (int)blockIdx.y + 1);  // a failing block, 1-based
        }
    };
    if (j < 0) {  // pivot 0: an empty dot
        if (wave == 0 && i == 0) {
            if (gl == 0) c.rng[0] = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN);
            const int code = ldl_col_pivot<W, G>(c, 0, gl, INT64_MAX, INT64_MIN, slots, signs, sign_mode,
                                                 B.sign_off, reg_eps, reg_delta, status);
            if (code) fail(code);
        }
        return;
    }
    if (j >= B.n) return;  // the whole workgroup: this block has no column j
    if (wave == 1) {
        if (lane == 0) {
            Val<W> dj, di;
            load(rec_at<W>((const uint32_t *)c.dstore, j), dj);
            recip(dj, di);
            store(dinv_sh, di);
            if (blockIdx.x == 0) store(rec_at<W>(c.dv, j), di);
        }
        __syncthreads();
        return;
    }
    if (wave == 2) {
        int64_t lo_w = INT64_MIN;
        const int64_t p = j + 1;
        if (G == 64 && blockIdx.x == 0 && p < B.n && j > 0 && c.first[p] < j) {
            const int4 rp = c.rng[p];  // row p's ranges over k < j
            int64_t hl, hh;
            range_hint<W>(rp.x, rp.y, rp.z, rp.w, hl, hh);
            if (hl <= hh && (hh - hl + 256) / 32 + 2 * W + 4 <= SLOTS) {
                lo_w = hl - 128;
                group_accumulate<W, ROW_THREADS, ROW_THREADS>(
                    c.first[p], j, lane, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(p)), 1,
                    rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(p)), 1, false, pre_win, SLOTS, lo_w);
            }
        }
        if (lane == 0) pre_lo = lo_w;
        __syncthreads();
        return;
    }
    const bool active = i < B.n;
    // row i's ranges so far (column 0 starts them)
    int4 ri = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN);
    Val<W> dj, dij, v, wv, diff, aij;
    bool dot_ok = true, zero_row = true;
    if (active) {
        if (j > 0) ri = c.rng[i];
        load(rec_at<W>((const uint32_t *)c.dstore, j), dj);
        zero_row = c.first[i] > j;
        if (!zero_row) {
            load(rec_at<W>(c.a, i + j * c.ld), aij);
            const int64_t lo = max(c.first[i], c.first[j]);
            const int4 rj = c.rng[j];
            int64_t hl, hh;
            range_hint<W>(ri.x, ri.y, rj.z, rj.w, hl, hh);
            Val<W> t;
            dot_ok = group_dot<W, G, ROW_THREADS>(lo, j, gl, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)),
                                                  1, rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j)), 1,
                                                  nullptr, false, slots, t, hl, hh);
            if (dot_ok) add_signed(aij, t, true, diff);
        }
    }
    __syncthreads();  // RN(1 / d_j) in dinv_sh, pivot j + 1's windows
    if (!active) return;
    if (!dot_ok) {
        fail(2);
        return;
    }
    if (zero_row) {
        set_zero(v);
    } else {
        load(dinv_sh, dij);
        mul(diff, dij, v);
    }
    mul(v, dj, wv);
    if (gl == (int)(j % G)) {  // the lane that reads term j back
        store(rec_at<W>(c.r, LdlCol<W>::tri(i) + j), v);
        store(rec_at<W>(c.w, LdlCol<W>::tri(i) + j), wv);
    }
    if (!is_zero(v)) {
        ri.x = min(ri.x, v.e);
        ri.y = max(ri.y, v.e);
    }
    if (!is_zero(wv)) {
        ri.z = min(ri.z, wv.e);
        ri.w = max(ri.w, wv.e);
    }
    if (gl == 0) c.rng[i] = ri;
    if (i != j + 1) return;
    int64_t hl, hh;
    range_hint<W>(ri.x, ri.y, ri.z, ri.w, hl, hh);
    const int64_t plo = pre_lo;
    if constexpr (G == 64) {
        if (plo != INT64_MIN) {  // pivot j + 1: the pre-accumulated windows plus the term v wv
            bool fits = true;
            if (!is_zero(v) && !is_zero(wv)) {
                const int64_t sc = (int64_t)v.e + wv.e - 64 * W;
                fits = sc >= plo && (sc - plo) / 32 + 2 * W + 4 <= SLOTS;
                if (fits && lane == 0) lds_add_product<W, ROW_THREADS>(pre_win, plo, v, wv, false);
            }
            if (fits) {
                Val<W> t;
                wave_finish<W, SLOTS, ROW_THREADS>(pre_win, lane, SLOTS, plo, t);
                const int code = ldl_col_pivot<W, G>(c, i, gl, hl, hh, slots, signs, sign_mode, B.sign_off,
                                                     reg_eps, reg_delta, status, &t);
                if (code) fail(code);
                return;
            }
        }
    }
    const int code = ldl_col_pivot<W, G>(c, i, gl, hl, hh, slots, signs, sign_mode, B.sign_off, reg_eps,
                                         reg_delta, status);
    if (code) fail(code);
}

// ldl_col_kernel with one row a wave and the column's chain spread over the
// wave (wavemp.h): the same three waves, launches, windows and roundings.
// Every operand a wave needs is loaded before the workgroup reads the
// status, so those loads overlap the launch's first barrier: the pivot
// record d_j (waves 0 and 1), the row's entry, ranges and dot terms
// k = lane, lane + 64 (wave 0; further terms are loaded in the dot), and
// the next pivot's a_jj, eps and delta (wave 0 of workgroup 0).  The dot
// runs over k in [0, j): the r and w entries left of a row's envelope are
// stored zeros, which the exact dot skips.
template <int W>
__global__ __launch_bounds__(3 * ROW_THREADS) void ldl_col_wave_kernel(
    uint32_t *A, uint32_t *dinv, const int32_t *first_all, const int8_t *signs, int sign_mode,
    const LdlBlock *blocks, uint32_t *scratch, const uint32_t *reg_eps, const uint32_t *reg_delta,
    int *status, int64_t j) {
    constexpr int SLOTS = Win<W>::SLOTS;
    extern __shared__ int64_t slots_sh[];  // [SLOTS][64] row windows, then [SLOTS][64] pivot windows
    __shared__ uint32_t dinv_sh[W + 2];
    __shared__ int stop;
    __shared__ int64_t pre_lo;
    __shared__ int64_t pre_tot[64];  // pivot j + 1's window totals, slot s at [s]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x == 0) stop = status[0] | status[2];  // a failed pivot, or an earlier row kernel's window overflow
    const LdlBlock B = blocks[blockIdx.y];
    const LdlCol<W> c(A, dinv, first_all, scratch, B);
    const int64_t i = j + 1 + blockIdx.x;
    int64_t *slots = slots_sh + lane;
    lds_i64 *pre_win = (lds_i64 *)(slots_sh + SLOTS * ROW_THREADS) + lane;
    // loads ahead of the status barrier (every address valid for 0 <= j < n)
    const bool live = j >= 0 && j < B.n;
    const bool row = live && wave == 0 && i < B.n;
    const bool next = row && i == j + 1;  // this wave forms pivot j + 1
    wmp::WV<W> dj = wmp::zero<W>(), aij = wmp::zero<W>(), ajj = wmp::zero<W>(), eps = wmp::zero<W>(),
               delta = wmp::zero<W>();
    Val<W> x0, y0, x1, y1;
    int64_t pre_t = 0, pre_b = INT64_MIN;
    uint32_t last_r = 0, last_w = 0;
    int4 ri = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN), rj = ri;
    int32_t fi = 0;
    int sg = 1;
    if (live && wave <= 1) dj = wmp::load<W>(rec_at<W>((const uint32_t *)c.dstore, j));
    if (row) {
        if (j > 0) ri = c.rng[i];
        rj = c.rng[j];
        fi = c.first[i];
        aij = wmp::load<W>(rec_at<W>((const uint32_t *)c.a, i + j * c.ld));
        const uint32_t *xr = rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i));
        const uint32_t *yr = rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j));
        if (lane < j) {
            load(rec_at<W>(xr, lane), x0);
            load(rec_at<W>(yr, lane), y0);
        } else {
            set_zero(x0);
            set_zero(y0);
        }
        if (lane + 64 < j) {
            load(rec_at<W>(xr, lane + 64), x1);
            load(rec_at<W>(yr, lane + 64), y1);
        } else {
            set_zero(x1);
            set_zero(y1);
        }
        if (j > 0) {  // the window launch j - 1 summed over k < j - 1, and the term k = j - 1
            const int64_t *pw = c.row_window(j, i);
            pre_t = lane < SLOTS ? pw[lane] : 0;
            pre_b = pw[SLOTS];
            last_r = wmp::load_word<W>(rec_at<W>(xr, j - 1));
            last_w = wmp::load_word<W>(rec_at<W>(yr, j - 1));
        }
        if (next) {
            ajj = wmp::load<W>(rec_at<W>((const uint32_t *)c.a, i + i * c.ld));
            if (reg_eps) {
                eps = wmp::load<W>(reg_eps);
                delta = wmp::load<W>(reg_delta);
            }
            sg = sign_mode == 0 ? 1 : (sign_mode == 1 ? signs[0] : signs[B.sign_off + i]);
        }
    }
    __syncthreads();
    if (stop != 0) return;  // the whole workgroup
    auto fail = [&](int code) {
        if (lane == 0) {
            atomicMax(status, code);
            if (code == 1) atomicMax(status + 3, (int)blockIdx.y + 1);  // a failing block, 1-based
        }
    };
    if (j < 0) {  // pivot 0: an empty dot
        if (wave == 0 && i == 0) {
            if (lane == 0) c.rng[0] = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN);
            const int code = ldl_col_pivot<W, 64>(c, 0, lane, INT64_MAX, INT64_MIN, slots, signs, sign_mode,
                                                  B.sign_off, reg_eps, reg_delta, status);
            if (code) fail(code);
        }
        return;
    }
    if (j >= B.n) return;  // the whole workgroup: this block has no column j
    if (wave == 1) {  // RN(1 / d_j) for this workgroup's rows; workgroup 0 stores it
        const wmp::WV<W> di = wmp::recip(dj);
        wmp::store(dinv_sh, di);
        if (blockIdx.x == 0) wmp::store(rec_at<W>(c.dv, j), di);
        __syncthreads();
        return;
    }
    if (wave == 2) {  // pivot j + 1's dot over k < j into the pivot windows
        int64_t lo_w = INT64_MIN;
        const int64_t p = j + 1;
        if (blockIdx.x == 0 && p < B.n && j > 0 && c.first[p] < j) {
            const int4 rp = c.rng[p];  // row p's ranges over k < j
            int64_t hl, hh;
            range_hint<W>(rp.x, rp.y, rp.z, rp.w, hl, hh);
            if (hl <= hh && (hh - hl + 256) / 32 + 2 * W + 4 <= SLOTS) {
                lo_w = hl - 128;
                group_accumulate<W, ROW_THREADS, ROW_THREADS>(
                    c.first[p], j, lane, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(p)), 1,
                    rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(p)), 1, false, pre_win, SLOTS, lo_w);
                pre_tot[lane] = window_total<ROW_THREADS>(pre_win, lane, SLOTS);  // summed off the chain
            }
        }
        if (lane == 0) pre_lo = lo_w;
        if (blockIdx.x > 0 && i < B.n && p < B.n) {
            // row i's dot for column p over k < j, for the next launch (which
            // adds the term k = j); the base leaves 128 bits of room each side
            int64_t lo_r = INT64_MIN;
            if (j > 0 && c.first[i] < j) {
                const int4 rp = c.rng[p], rr = c.rng[i];  // w of row p, r of row i, over k < j (or wider)
                int64_t hl, hh;
                range_hint<W>(rr.x, rr.y, rp.z, rp.w, hl, hh);
                if (hl <= hh && (hh - hl + 256) / 32 + 2 * W + 4 <= SLOTS) {
                    lo_r = hl - 128;
                    group_accumulate<W, ROW_THREADS, ROW_THREADS>(
                        max(c.first[i], c.first[p]), j, lane, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)), 1,
                        rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(p)), 1, false, pre_win, SLOTS, lo_r);
                    const int64_t T = window_total<ROW_THREADS>(pre_win, lane, SLOTS);
                    if (lane < SLOTS) c.row_window(p, i)[lane] = T;
                }
            }
            if (lane == 0) c.row_window(p, i)[SLOTS] = lo_r;
        }
        __syncthreads();
        return;
    }
    // wave 0: row i's entry (i, j)
    wmp::WV<W> diff = wmp::zero<W>();
    bool dot_ok = true;
    const bool zero_row = !row || fi > j;
    if (!zero_row) {
        int64_t hl, hh;
        range_hint<W>(ri.x, ri.y, rj.z, rj.w, hl, hh);
        wmp::WV<W> t;
        bool done = false;
        if (pre_b != INT64_MIN) {  // the window summed by launch j - 1 plus the term k = j - 1
            const wmp::WV<W> rl = wmp::from_word<W>(last_r), wl = wmp::from_word<W>(last_w);
            bool fits = true;
            if (!wmp::is_zero(rl) && !wmp::is_zero(wl)) {
                const int64_t sc = (int64_t)rl.e + wl.e - 64 * W;
                fits = sc >= pre_b && (sc - pre_b) / 32 + 2 * W + 4 <= SLOTS;
                if (fits) pre_t = wmp::add_product<W, SLOTS>(pre_t, rl, wl, sc - pre_b, false);
            }
            if (fits) {
                t = wmp::round_window<W, SLOTS>(pre_t, pre_b);
                done = true;
            }
        }
        if (done) {
        } else if (hl <= hh && (hh - hl) / 32 + 2 * W + 4 <= SLOTS) {  // the window from the ranges; terms preloaded
            const int used = (int)min((int64_t)SLOTS, (hh - hl) / 32 + 2 * W + 1);
            lds_i64 *ls = (lds_i64 *)slots;
            for (int sl = 0; sl < used; sl++) ls[sl * ROW_THREADS] = 0;
            if (!is_zero(x0) && !is_zero(y0)) lds_add_product<W, ROW_THREADS>(ls, hl, x0, y0, false);
            if (!is_zero(x1) && !is_zero(y1)) lds_add_product<W, ROW_THREADS>(ls, hl, x1, y1, false);
            for (int64_t k = lane + 128; k < j; k += 64) {
                Val<W> x, y;
                load(rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i) + k), x);
                load(rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j) + k), y);
                if (!is_zero(x) && !is_zero(y)) lds_add_product<W, ROW_THREADS>(ls, hl, x, y, false);
            }
            t = wmp::round_window<W, SLOTS>(window_total<ROW_THREADS>(ls, lane, used), hl);
        } else {
            dot_ok = wave_dot<W, ROW_THREADS>(0, j, lane, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)), 1,
                                              rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j)), 1, slots, t,
                                              hl, hh);
        }
        if (dot_ok) diff = wmp::add(aij, t, true);
    }
    __syncthreads();  // RN(1 / d_j) in dinv_sh, pivot j + 1's windows
    if (!row) return;
    if (!dot_ok) {
        fail(2);
        return;
    }
    const wmp::WV<W> v = zero_row ? wmp::zero<W>() : wmp::mul(diff, wmp::load<W>(dinv_sh));
    const wmp::WV<W> wv = wmp::mul(v, dj);
    wmp::store(rec_at<W>(c.r, LdlCol<W>::tri(i) + j), v);
    wmp::store(rec_at<W>(c.w, LdlCol<W>::tri(i) + j), wv);
    if (!wmp::is_zero(v)) {
        ri.x = min(ri.x, v.e);
        ri.y = max(ri.y, v.e);
    }
    if (!wmp::is_zero(wv)) {
        ri.z = min(ri.z, wv.e);
        ri.w = max(ri.w, wv.e);
    }
    if (lane == 0) c.rng[i] = ri;
    if (!next) return;
    const int64_t plo = pre_lo;
    if (plo != INT64_MIN) {  // pivot j + 1: the pre-accumulated windows plus the term v wv
        int64_t T = pre_tot[lane];
        bool fits = true;
        if (!wmp::is_zero(v) && !wmp::is_zero(wv)) {
            const int64_t sc = (int64_t)v.e + wv.e - 64 * W;
            fits = sc >= plo && (sc - plo) / 32 + 2 * W + 4 <= SLOTS;
            if (fits) T = wmp::add_product<W, SLOTS>(T, v, wv, sc - plo, false);
        }
        if (fits) {
            const int code = ldl_col_pivot_wave<W>(c, i, wmp::round_window<W, SLOTS>(T, plo), ajj, reg_eps != nullptr,
                                                   eps, delta, sg, status);
            if (code != 3) {
                if (code) fail(code);
                return;
            }
        }
    }
    int64_t hl, hh;
    range_hint<W>(ri.x, ri.y, ri.z, ri.w, hl, hh);
    __threadfence_block();  // the whole dot reads the entries this wave's lanes just stored
    const int code = ldl_col_pivot<W, 64>(c, i, lane, hl, hh, slots, signs, sign_mode, B.sign_off, reg_eps,
                                          reg_delta, status);
    if (code) fail(code);
}

// ---------------------------------------------------------------------------
// ldl_dots in one launch: workgroup (x, y) owns row i = x of block y and forms
// its entries (i, j), j < i, column by column, then (row i) pivot i.  The
// column order that ldl_col_wave_kernel gets from kernel boundaries comes
// from flags in the block's scratch: row j sets pivot_ready[j] once d_j, its
// r and w rows and its ranges are stored (agent-scope release), and every
// entry (i, j) waits for it (acquire); row i sets progress[i] = j + 1 after
// storing entry (i, j), which the look-ahead of later rows waits for (their
// column j' + 1 dot over k < j' reads row j' + 1).  A row waits only on rows
// of lower index, so in-order dispatch drains the grid even when not every
// workgroup is resident.  Every wait is bounded (status code 4 after
// LDL_WAIT_LIMIT ticks of the 100 MHz clock) and ends on any status code, and
// a workgroup's three waves leave together after a barrier, so every wave
// reaches the end.  Per step: wave 1 forms RN(1 / d_j) while wave 0 finishes
// the entry's dot (the window wave 2 summed in the previous step plus the
// term k = j - 1) and wave 2 sums the next column's window (or, in row i's
// last step, pivot i's window over its own row); after the barrier wave 0
// forms the entry and its W value, and in the last step pivot i.  Same
// arithmetic, bits and status codes as the column launches.
constexpr uint64_t LDL_WAIT_LIMIT = 200000000ull;  // 2 s of the 100 MHz clock

__device__ __forceinline__ bool ldl_wait(const int *flag, int target, const int *status) {
    // The poll is a relaxed load; one agent-scope acquire fence after the
    // flag is seen orders the later loads (an acquire load per poll would
    // invalidate the L2 of this XCD on every pass).
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
             __hip_atomic_load(status + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
            return false;
        if (__builtin_amdgcn_s_memrealtime() - t0 > LDL_WAIT_LIMIT) {
            if (wmp::lane() == 0) atomicMax((int *)status, 4);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// flag = value after this wave's earlier stores (agent-scope release)
__device__ __forceinline__ void ldl_publish(int *flag, int value) {
    if (wmp::lane() == 0) __hip_atomic_store(flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int W>
__global__ void ldl_flags_clear_kernel(uint32_t *A, uint32_t *dinv, const int32_t *first_all, uint32_t *scratch,
                                       const LdlBlock *blocks) {
    const LdlBlock B = blocks[blockIdx.y];
    const LdlCol<W> c(A, dinv, first_all, scratch, B);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * B.n; e += (int64_t)gridDim.x * blockDim.x)
        c.flags[e] = 0;
}

template <int W>
__global__ __launch_bounds__(3 * ROW_THREADS) void ldl_rows_kernel(
    uint32_t *A, uint32_t *dinv, const int32_t *first_all, const int8_t *signs, int sign_mode,
    const LdlBlock *blocks, int nblocks, uint32_t *scratch, const uint32_t *reg_eps, const uint32_t *reg_delta,
    int *status) {
    constexpr int SLOTS = Win<W>::SLOTS;
    extern __shared__ int64_t slots_sh[];  // [SLOTS][64] row windows, then [SLOTS][64] look-ahead windows
    __shared__ uint32_t dinv_sh[W + 2];
    __shared__ int64_t nxt_tot[2][64];  // row i's window for column j (slot s at [s]), by column parity
    __shared__ int64_t nxt_base[2];
    __shared__ int64_t pre_tot[64];     // pivot i's window over k < i - 1
    __shared__ int64_t pre_lo;
    __shared__ int ok_sh[3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // row-major over the blocks: workgroup (row i, block bk) at linear index
    // i * nblocks + bk, so every block's rows start together and each row's
    // waits (on lower rows of its block) are on earlier workgroups
    const int bk = (int)(blockIdx.x % nblocks);
    const int64_t i = blockIdx.x / nblocks;
    const LdlBlock B = blocks[bk];
    if (i >= B.n) return;  // the whole workgroup
    const LdlCol<W> c(A, dinv, first_all, scratch, B);
    int *const pivot_ready = c.flags, *const progress = c.flags + c.n;
    int64_t *slots = slots_sh + lane;
    lds_i64 *pre_win = (lds_i64 *)(slots_sh + SLOTS * ROW_THREADS) + lane;
    auto fail = [&](int code) {
        if (lane == 0) {
            atomicMax(status, code);
            if (code == 1) atomicMax(status + 3, bk + 1);  // a failing block, 1-based
        }
    };
    if (i == 0) {  // pivot 0: an empty dot
        if (wave == 0) {
            if (lane == 0) c.rng[0] = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN);
            const int code = ldl_col_pivot<W, 64>(c, 0, lane, INT64_MAX, INT64_MIN, slots, signs, sign_mode,
                                                  B.sign_off, reg_eps, reg_delta, status);
            if (code)
                fail(code);
            else
                ldl_publish(pivot_ready, 1);
        }
        return;  // the whole workgroup (no barrier follows)
    }
    const int64_t fi = c.first[i];
    if (threadIdx.x < 2) nxt_base[threadIdx.x] = INT64_MIN;
    // the pivot chain (wave 0: entries and pivot, wave 1: reciprocals) ahead
    // of the look-ahead wave where workgroups share a SIMD
    if (g_chain_prio && wave < 2) __builtin_amdgcn_s_setprio(2);
    __syncthreads();
    // wave 0: pivot i's operands, loaded now so that no global load sits on
    // the pivot's critical path
    wmp::WV<W> aii_w = wmp::zero<W>(), eps_w = wmp::zero<W>(), delta_w = wmp::zero<W>();
    if (wave == 0) {
        aii_w = wmp::load<W>(rec_at<W>((const uint32_t *)c.a, i + i * c.ld));
        if (reg_eps) {
            eps_w = wmp::load<W>(reg_eps);
            delta_w = wmp::load<W>(reg_delta);
        }
    }
    const int sg_i = sign_mode == 0 ? 1 : (sign_mode == 1 ? signs[0] : signs[B.sign_off + i]);
    int4 ri = make_int4(INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN);  // wave 0: row i's ranges so far
    wmp::WV<W> vprev = wmp::zero<W>();                                  // wave 0: r_{i, j - 1}
    for (int64_t j = 0; j < i; j++) {
        const int par = (int)(j & 1);
        const bool last = j == i - 1;  // this step also forms pivot i
        const bool zero_row = fi > j;  // no entry (i, j): left of the envelope
        bool okw = true;
        wmp::WV<W> dj = wmp::zero<W>(), diff = wmp::zero<W>();
        bool dot_ok = true;
        if (wave == 1) {  // RN(1 / d_j); row j + 1 also stores it
            okw = ldl_wait(pivot_ready + j, 1, status);
            if (okw) {
                const wmp::WV<W> di = wmp::recip(wmp::load<W>(rec_at<W>((const uint32_t *)c.dstore, j)));
                wmp::store(dinv_sh, di);
                if (last) wmp::store(rec_at<W>(c.dv, j), di);
            }
        } else if (wave == 2) {
            if (last) {  // pivot i's dot over k < j (its own r and w rows)
                int64_t lo_w = INT64_MIN;
                if (j > 0 && fi < j) {
                    const int4 rp = c.rng[i];  // row i's ranges over k < j
                    int64_t hl, hh;
                    range_hint<W>(rp.x, rp.y, rp.z, rp.w, hl, hh);
                    if (hl <= hh && (hh - hl + 256) / 32 + 2 * W + 4 <= SLOTS) {
                        lo_w = hl - 128;
                        group_accumulate<W, ROW_THREADS, ROW_THREADS>(
                            fi, j, lane, rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)), 1,
                            rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(i)), 1, false, pre_win, SLOTS, lo_w);
                        pre_tot[lane] = window_total<ROW_THREADS>(pre_win, lane, SLOTS);
                    }
                }
                if (lane == 0) pre_lo = lo_w;
            } else {  // row i's dot for column p = j + 1 over k < j (row p through column j - 1)
                const int64_t p = j + 1;
                int64_t lo_r = INT64_MIN;
                if (j > 0 && fi < j) {
                    okw = ldl_wait(progress + p, (int)j, status);
                    if (okw) {
                        const int4 rp = c.rng[p], rr = c.rng[i];  // w of row p, r of row i, over k < j (or wider)
                        int64_t hl, hh;
                        range_hint<W>(rr.x, rr.y, rp.z, rp.w, hl, hh);
                        if (hl <= hh && (hh - hl + 256) / 32 + 2 * W + 4 <= SLOTS) {
                            lo_r = hl - 128;
                            group_accumulate<W, ROW_THREADS, ROW_THREADS>(
                                max(fi, (int64_t)c.first[p]), j, lane,
                                rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)), 1,
                                rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(p)), 1, false, pre_win, SLOTS, lo_r);
                            const int64_t T = window_total<ROW_THREADS>(pre_win, lane, SLOTS);
                            nxt_tot[par ^ 1][lane] = lane < SLOTS ? T : 0;
                        }
                    }
                }
                if (lane == 0) nxt_base[par ^ 1] = lo_r;
            }
        } else {  // wave 0: entry (i, j) up to its product with RN(1 / d_j)
            // a_ij is input: its load runs beside the wait
            const wmp::WV<W> aij = zero_row ? wmp::zero<W>() : wmp::load<W>(rec_at<W>((const uint32_t *)c.a, i + j * c.ld));
            okw = ldl_wait(pivot_ready + j, 1, status);
            if (okw) {
                dj = wmp::load<W>(rec_at<W>((const uint32_t *)c.dstore, j));
                if (!zero_row) {
                    const int4 rj = c.rng[j];
                    int64_t hl, hh;
                    range_hint<W>(ri.x, ri.y, rj.z, rj.w, hl, hh);
                    wmp::WV<W> t;
                    bool done = false;
                    const int64_t pre_b = nxt_base[par];
                    if (pre_b != INT64_MIN) {  // the window wave 2 summed over k < j - 1, and the term k = j - 1
                        int64_t pre_t = lane < SLOTS ? nxt_tot[par][lane] : 0;
                        const wmp::WV<W> wl =
                            wmp::load<W>(rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j) + j - 1));
                        bool fits = true;
                        if (!wmp::is_zero(vprev) && !wmp::is_zero(wl)) {
                            const int64_t sc = (int64_t)vprev.e + wl.e - 64 * W;
                            fits = sc >= pre_b && (sc - pre_b) / 32 + 2 * W + 4 <= SLOTS;
                            if (fits) pre_t = wmp::add_product<W, SLOTS>(pre_t, vprev, wl, sc - pre_b, false);
                        }
                        if (fits) {
                            t = wmp::round_window<W, SLOTS>(pre_t, pre_b);
                            done = true;
                        }
                    }
                    if (done) {
                    } else if (hl <= hh && (hh - hl) / 32 + 2 * W + 4 <= SLOTS) {  // the window from the ranges
                        const int used = (int)min((int64_t)SLOTS, (hh - hl) / 32 + 2 * W + 1);
                        lds_i64 *ls = (lds_i64 *)slots;
                        for (int sl = 0; sl < used; sl++) ls[sl * ROW_THREADS] = 0;
                        for (int64_t k = lane; k < j; k += 64) {
                            Val<W> x, y;
                            load(rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i) + k), x);
                            load(rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j) + k), y);
                            if (!is_zero(x) && !is_zero(y)) lds_add_product<W, ROW_THREADS>(ls, hl, x, y, false);
                        }
                        t = wmp::round_window<W, SLOTS>(window_total<ROW_THREADS>(ls, lane, used), hl);
                    } else if (j == 0) {
                        t = wmp::zero<W>();
                    } else {
                        dot_ok = wave_dot<W, ROW_THREADS>(0, j, lane,
                                                          rec_at<W>((const uint32_t *)c.r, LdlCol<W>::tri(i)), 1,
                                                          rec_at<W>((const uint32_t *)c.w, LdlCol<W>::tri(j)), 1,
                                                          slots, t, hl, hh);
                    }
                    if (dot_ok) diff = wmp::add(aij, t, true);
                }
            }
        }
        if (lane == 0) ok_sh[wave] = okw ? 1 : 0;
        __syncthreads();  // RN(1 / d_j) in dinv_sh, the windows of wave 2
        const bool go = ok_sh[0] && ok_sh[1] && ok_sh[2];
        if (go && wave == 0) {
            if (!dot_ok) {
                fail(2);
            } else {
                const wmp::WV<W> v = zero_row ? wmp::zero<W>() : wmp::mul(diff, wmp::load<W>(dinv_sh));
                const wmp::WV<W> wv = wmp::mul(v, dj);
                wmp::store(rec_at<W>(c.r, LdlCol<W>::tri(i) + j), v);
                wmp::store(rec_at<W>(c.w, LdlCol<W>::tri(i) + j), wv);
                if (!wmp::is_zero(v)) {
                    ri.x = min(ri.x, v.e);
                    ri.y = max(ri.y, v.e);
                }
                if (!wmp::is_zero(wv)) {
                    ri.z = min(ri.z, wv.e);
                    ri.w = max(ri.w, wv.e);
                }
                if (lane == 0) c.rng[i] = ri;
                vprev = v;
                if (!last) {
                    ldl_publish(progress + i, (int)(j + 1));
                } else {  // pivot i: the window of wave 2 plus the term v wv
                    int code = 3;
                    const int64_t plo = pre_lo;
                    if (plo != INT64_MIN) {
                        const wmp::WV<W> &aii = aii_w, &eps = eps_w, &delta = delta_w;
                        const int sg = sg_i;
                        int64_t T = pre_tot[lane];
                        bool fits = true;
                        if (!wmp::is_zero(v) && !wmp::is_zero(wv)) {
                            const int64_t sc = (int64_t)v.e + wv.e - 64 * W;
                            fits = sc >= plo && (sc - plo) / 32 + 2 * W + 4 <= SLOTS;
                            if (fits) T = wmp::add_product<W, SLOTS>(T, v, wv, sc - plo, false);
                        }
                        if (fits)
                            code = ldl_col_pivot_wave<W>(c, i, wmp::round_window<W, SLOTS>(T, plo), aii,
                                                         reg_eps != nullptr, eps, delta, sg, status);
                    }
                    if (code == 3) {  // the whole dot (also when the pivot cancels: rounded once exactly)
                        int64_t hl, hh;
                        range_hint<W>(ri.x, ri.y, ri.z, ri.w, hl, hh);
                        __threadfence_block();  // the whole dot reads the entries this wave's lanes just stored
                        code = ldl_col_pivot<W, 64>(c, i, lane, hl, hh, slots, signs, sign_mode, B.sign_off,
                                                    reg_eps, reg_delta, status);
                    }
                    if (code)
                        fail(code);
                    else
                        ldl_publish(pivot_ready + i, 1);
                }
            }
        }
        __syncthreads();  // the step's stores before the next step's look-ahead reads them
        if (!go) break;   // the whole workgroup
    }
}

// the unit lower factor of every block (after the last column launch)
template <int W>
__global__ void ldl_col_writeout_kernel(uint32_t *A, uint32_t *dinv, const LdlBlock *blocks, uint32_t *scratch,
                                        const int *status) {
    if (status[0] != 0) return;
    const LdlBlock B = blocks[blockIdx.y];
    const int64_t n = B.n, e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0 && n > 0) {  // RN(1 / d) of the last pivot (no column launch follows it)
        const uint32_t *dstore = scratch + B.scratch_off * (W + 2) + n * (n + 1) * (W + 2);
        Val<W> d, di;
        load(rec_at<W>(dstore, n - 1), d);
        recip(d, di);
        store(rec_at<W>(dinv + B.dinv_off * (W + 2), n - 1), di);
    }
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    if (i < j) return;
    uint32_t *a = A + B.a_off * (W + 2);
    const uint32_t *r = scratch + B.scratch_off * (W + 2);
    Val<W> v;
    if (i == j) {
        for (int u = 0; u < W; u++) v.m[u] = 0;
        v.m[W - 1] = 0x80000000u;
        v.kind = mpd::KIND_REG;
        v.e = 1;
    } else {
        load(rec_at<W>(r, i * (i + 1) / 2 + j), v);
    }
    store(rec_at<W>(a, i + j * B.ld), v);
}

// ---------------------------------------------------------------------------
// exact-dot Cholesky, triangular solve and solve (mpfr.rs:726-872,
// traits.rs:85-113)
// ---------------------------------------------------------------------------
// xpotrf, one launch per column j (j = -1 forms the first pivot only):
// status[0] = info, status[2] |= 1 on a dot beyond its window; every launch
// returns at once when either is set.  One lane group per row i > j forms
// L_ij = RN(RN(a_ij - RN(sum_{k<j} L_ik L_jk)) / d_j); the group of row j+1
// then forms the next pivot d = RN(a - RN(sum_{k<=j} L_{j+1,k}^2)), stores
// it, checks it (finite and positive) and stores RN(sqrt(d)).  Entry (i, k)
// of the stored triangle is a[i + k lda] (lower) or a[k + i lda] (upper).
template <int W, int G>
__global__ __launch_bounds__(ROW_THREADS) void potrf_col_kernel(uint32_t *a, int64_t n, int64_t lda,
                                                                int lower, int64_t j, int *status) {
    extern __shared__ int64_t slots_sh[];
    const int gl = threadIdx.x & (G - 1);
    const int64_t i = j + 1 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (i >= n || status[0] != 0 || status[2] != 0) return;
    int64_t *slots = slots_sh + threadIdx.x;
    const int64_t rs = lower ? lda : 1;  // record stride along a row of the triangle
    auto row = [&](int64_t r) { return rec_at<W>((const uint32_t *)a, lower ? r : r * lda); };
    if (j >= 0) {
        Val<W> t, aij, diff, dj, v;
        if (!group_dot<W, G, ROW_THREADS>(0, j, gl, row(i), rs, row(j), rs, nullptr, false, slots, t,
                                          INT64_MAX, INT64_MIN)) {
            if (gl == 0) atomicOr(status + 2, 1);
            return;
        }
        uint32_t *q = rec_at<W>(a, lower ? i + j * lda : j + i * lda);
        load(q, aij);
        load(rec_at<W>((const uint32_t *)a, j + j * lda), dj);
        add_signed(aij, t, true, diff);
        div(diff, dj, v);
        if (gl == (int)(j % G)) store(q, v);  // the lane that reads term j below
    }
    if (i != j + 1) return;
    Val<W> t, aii, d;
    if (!group_dot<W, G, ROW_THREADS>(0, i, gl, row(i), rs, row(i), rs, nullptr, false, slots, t, INT64_MAX,
                                      INT64_MIN)) {
        if (gl == 0) atomicOr(status + 2, 1);
        return;
    }
    if (gl != 0) return;
    uint32_t *pii = rec_at<W>(a, i + i * lda);
    load(pii, aii);
    add_signed(aii, t, true, d);
    store(pii, d);
    if (!is_regular(d) || d.kind < 0) {
        status[0] = (int)(i + 1);
        return;
    }
    Val<W> r;
    sqrt(d, r);
    store(pii, r);
}

// xtrsm_lower: per right-hand side (one lane group each), x_i =
// RN(RN(b_i - RN(sum_k L_ik x_k)) / L_ii), ascending, or with L^T
// (transpose: terms L_ki, k > i) descending; a is n x n (ld n), b n x nrhs.
template <int W, int G>
__global__ __launch_bounds__(ROW_THREADS) void trsm_lower_kernel(const uint32_t *a, int64_t n, uint32_t *b,
                                                                 int64_t nrhs, int transpose, int *fail) {
    extern __shared__ int64_t slots_sh[];
    const int gl = threadIdx.x & (G - 1);
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (c >= nrhs) return;
    int64_t *slots = slots_sh + threadIdx.x;
    uint32_t *col = rec_at<W>(b, c * n);
    for (int64_t step = 0; step < n; step++) {
        const int64_t i = transpose ? n - 1 - step : step;
        Val<W> t, bi, diff, aii, x;
        const bool ok =
            transpose ? group_dot<W, G, ROW_THREADS>(i + 1, n, gl, rec_at<W>(a, i * n), 1, col, 1, nullptr,
                                                     false, slots, t, INT64_MAX, INT64_MIN)
                      : group_dot<W, G, ROW_THREADS>(0, i, gl, rec_at<W>(a, i), n, col, 1, nullptr, false,
                                                     slots, t, INT64_MAX, INT64_MIN);
        if (!ok) {
            if (gl == 0) atomicOr(fail, 1);
            return;
        }
        load(rec_at<W>(col, i), bi);
        load(rec_at<W>(a, i + i * n), aii);
        add_signed(bi, t, true, diff);
        div(diff, aii, x);
        if (gl == (int)(i % G)) store(rec_at<W>(col, i), x);
    }
}

// xpotrs: one lane per right-hand side runs the forward and backward chains
// v = RN(v - a_ik x_k) (one rounding, the fused multiply-add) and
// x_i = RN(v / a_ii) in the reference's order.
template <int W>
__global__ __launch_bounds__(ROW_THREADS) void potrs_kernel(const uint32_t *a, int64_t n, int64_t lda,
                                                            int lower, uint32_t *b, int64_t ldb,
                                                            int64_t nrhs) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nrhs) return;
    uint32_t *col = rec_at<W>(b, c * ldb);
    auto entry = [&](int64_t i, int64_t k) { return rec_at<W>(a, lower ? i + k * lda : k + i * lda); };
    for (int pass = 0; pass < 2; pass++) {
        for (int64_t step = 0; step < n; step++) {
            const int64_t i = pass == 0 ? step : n - 1 - step;
            Val<W> v, w, ak, xk;
            load(rec_at<W>(col, i), v);
            const int64_t k0 = pass == 0 ? 0 : i + 1, k1 = pass == 0 ? i : n;
            for (int64_t k = k0; k < k1; k++) {
                // forward reads L_ik, backward L_ki (the transposed factor)
                load(pass == 0 ? entry(i, k) : entry(k, i), ak);
                load(rec_at<W>(col, k), xk);
                fms(v, ak, xk, w);
                v = w;
            }
            Val<W> aii, x;
            load(rec_at<W>(a, i + i * lda), aii);
            div(v, aii, x);
            store(rec_at<W>(col, i), x);
        }
    }
}

// ---------------------------------------------------------------------------
// FMA sweeps of the arrow solves (DenseLeaf::forward_many / backward,
// arrow.rs:1155-1183): one workgroup per (factor, right-hand side), row
// i + r T of the column in thread i's register r.  Every entry sees the
// reference's sequence of single-rounding updates x_i = RN(x_i - l x_k):
// the forward sweep applies k = 0, 1, ... (right-looking: step k broadcasts
// the final x_k through LDS and every later row applies it, which keeps
// each row's ascending order); the backward sweep takes blocks of `back`
// rows from the bottom, and a row applies a block's columns in ascending
// order: the block's own rows finish one by one from its last (each over
// the block's later rows), then every earlier row applies the whole block
// (DenseLeaf::backward_blocked, arrow.rs:1104-1140).
struct SweepDesc {  // one factor and its right-hand sides
    int64_t l_off, n, x_off, ldx, d_off;
};

template <int W, int R>
__global__ __launch_bounds__(1024) void arrow_forward_kernel(const uint32_t *L_all, const uint32_t *dinv_all,
                                                            const SweepDesc *descs, uint32_t *X, uint32_t *Wout,
                                                            uint32_t *Vout, int64_t ldw) {
    __shared__ uint32_t xk_sh[2][W + 2];
    const SweepDesc d = descs[blockIdx.y];
    const int64_t c = blockIdx.x, n = d.n, T = blockDim.x;
    const uint32_t *L = L_all + d.l_off * (W + 2);
    uint32_t *x = X + (d.x_off + c * d.ldx) * (W + 2);
    Val<W> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t i = threadIdx.x + r * T;
        if (i < n) load(rec_at<W>((const uint32_t *)x, i), v[r]);
        else set_zero(v[r]);
    }
    for (int64_t k = 0; k < n; k++) {
        const int par = (int)(k & 1);
        if (threadIdx.x == k % T) {  // x_k is final: every update k' < k is in
#pragma unroll
            for (int r = 0; r < R; r++)
                if (r == k / T) store(xk_sh[par], v[r]);
        }
        __syncthreads();
        Val<W> xk;
        load((const uint32_t *)xk_sh[par], xk);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t i = threadIdx.x + r * T;
            if (i > k && i < n) {
                Val<W> l, t;
                load(rec_at<W>(L, i + k * n), l);
                fms(v[r], l, xk, t);
                v[r] = t;
            }
        }
    }
    // w (the forward result) and, when dinv is given, v = RN(w D^-1)
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t i = threadIdx.x + r * T;
        if (i >= n) continue;
        if (Wout) store(rec_at<W>(Wout, d.x_off + i + c * ldw), v[r]);
        if (dinv_all) {
            Val<W> di, sv;
            load(rec_at<W>(dinv_all, d.d_off + i), di);
            mul(v[r], di, sv);
            if (Vout) store(rec_at<W>(Vout, d.x_off + i + c * ldw), sv);
            else store(rec_at<W>(x, i), sv);
        } else if (!Wout) {
            store(rec_at<W>(x, i), v[r]);
        }
    }
}

template <int W, int R>
__global__ __launch_bounds__(1024) void arrow_backward_kernel(const uint32_t *L_all, const SweepDesc *descs,
                                                             uint32_t *X, int back, int wave_chain) {
    extern __shared__ uint32_t blk_sh[];  // [back][W + 2]: the current block's final x
    const SweepDesc d = descs[blockIdx.y];
    const int64_t c = blockIdx.x, n = d.n, T = blockDim.x;
    const uint32_t *L = L_all + d.l_off * (W + 2);
    uint32_t *x = X + (d.x_off + c * d.ldx) * (W + 2);
    Val<W> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t i = threadIdx.x + r * T;
        if (i < n) load(rec_at<W>((const uint32_t *)x, i), v[r]);
        else set_zero(v[r]);
    }
    for (int64_t r1 = n; r1 > 0;) {
        const int64_t r0 = r1 > back ? r1 - back : 0;
        // the block's rows, last first: row i applies k = i + 1 .. r1 - 1
        if (wave_chain) {  // on wave 0 (wave fms), the rows' values through LDS
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int64_t i = threadIdx.x + r * T;
                if (i >= r0 && i < r1) store(blk_sh + (i - r0) * (W + 2), v[r]);
            }
            __syncthreads();
            if (threadIdx.x < 64) {
                for (int64_t i = r1 - 1; i >= r0; i--) {
                    wmp::WV<W> vw = wmp::load<W>(blk_sh + (i - r0) * (W + 2));
                    uint32_t lw = i + 1 < r1 ? wmp::load_word<W>(rec_at<W>(L, (i + 1) + i * n)) : 0u;
                    for (int64_t k = i + 1; k < r1; k++) {
                        const wmp::WV<W> l = wmp::from_word<W>(lw);
                        if (k + 1 < r1) lw = wmp::load_word<W>(rec_at<W>(L, (k + 1) + i * n));
                        vw = wmp::fms<W>(vw, l, wmp::load<W>(blk_sh + (k - r0) * (W + 2)));
                    }
                    wmp::store(blk_sh + (i - r0) * (W + 2), vw);
                    __threadfence_block();  // the next rows read this record's words from other lanes
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int64_t i = threadIdx.x + r * T;
                if (i >= r0 && i < r1) load((const uint32_t *)(blk_sh + (i - r0) * (W + 2)), v[r]);
            }
        } else
        for (int64_t i = r1 - 1; i >= r0; i--) {
            if (threadIdx.x == i % T) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (r != i / T) continue;
                    for (int64_t k = i + 1; k < r1; k++) {
                        Val<W> l, xk, t;
                        load(rec_at<W>(L, k + i * n), l);
                        load((const uint32_t *)(blk_sh + (k - r0) * (W + 2)), xk);
                        fms(v[r], l, xk, t);
                        v[r] = t;
                    }
                    store(blk_sh + (i - r0) * (W + 2), v[r]);
                }
            }
            __syncthreads();
        }
        // every earlier row applies the block, ascending
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t i = threadIdx.x + r * T;
            if (i >= r0) continue;
            for (int64_t k = r0; k < r1; k++) {
                Val<W> l, xk, t;
                load(rec_at<W>(L, k + i * n), l);
                load((const uint32_t *)(blk_sh + (k - r0) * (W + 2)), xk);
                fms(v[r], l, xk, t);
                v[r] = t;
            }
        }
        __syncthreads();  // the block's LDS rows are read before the next block writes them
        r1 = r0;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t i = threadIdx.x + r * T;
        if (i < n) store(rec_at<W>(x, i), v[r]);
    }
}

// ---------------------------------------------------------------------------
// Blocked FMA sweeps of one large factor (the border's): the same sequence
// of single-rounding updates as the two kernels above, with the rows after
// (before) a block spread over many workgroups.  Per block of b <= 32 rows:
// a diagonal kernel finishes the block's own rows on one wave per
// right-hand side (the block's strictly lower triangle of L staged in LDS
// first), then an update kernel applies the block's b columns, ascending,
// to every later (earlier) row, one thread a row, L's entries fetched four
// steps ahead.  Forward: row i takes k ascending, whatever the blocks.
// Backward (blocks of `back` rows from the bottom): the block's rows finish
// from its last, each over the block's later rows ascending; then every
// earlier row applies the whole block ascending.
// ---------------------------------------------------------------------------
constexpr int SWEEP_BLOCK = 32;

// the block's strictly lower triangle, entry (i, k) (k < i, both relative to
// r0) at tri(i) + k, into LDS
template <int W>
__device__ __forceinline__ void stage_block_tri(const uint32_t *L, int64_t n, int64_t r0, int b, uint32_t *sh) {
    const int total = b * (b - 1) / 2;
    for (int e = threadIdx.x; e < total * (W + 2); e += blockDim.x) {
        const int rec = e / (W + 2), word = e % (W + 2);
        int i = (int)((sqrtf(8.0f * rec + 1.0f) + 1.0f) * 0.5f);
        while (i * (i - 1) / 2 > rec) i--;
        while ((i + 1) * i / 2 <= rec) i++;
        const int k = rec - i * (i - 1) / 2;
        sh[e] = rec_at<W>(L, r0 + i + (r0 + k) * n)[word];
    }
}

template <int W>
__global__ __launch_bounds__(64) void sweep_fwd_diag_kernel(const uint32_t *L, int64_t n, uint32_t *X, int64_t ldx,
                                                            int64_t k0, int64_t k1) {
    extern __shared__ uint32_t tri_sh[];
    __shared__ uint32_t xk_sh[2][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.x * ldx * (W + 2);
    const int b = (int)(k1 - k0), lane = threadIdx.x;
    stage_block_tri<W>(L, n, k0, b, tri_sh);
    const bool mine = lane < b;
    Val<W> v;
    if (mine) load(rec_at<W>((const uint32_t *)x, k0 + lane), v);
    __syncthreads();
    for (int k = 0; k < b; k++) {
        const int par = k & 1;
        if (lane == k) store(xk_sh[par], v);  // final: every k' < k applied
        __syncthreads();
        if (mine && lane > k) {
            Val<W> xk, l, t;
            load((const uint32_t *)xk_sh[par], xk);
            load((const uint32_t *)(tri_sh + (lane * (lane - 1) / 2 + k) * (W + 2)), l);
            fms(v, l, xk, t);
            v = t;
        }
    }
    if (mine) store(rec_at<W>(x, k0 + lane), v);
}

// The same forward block on 16 waves, each holding two of the block's rows
// in wave form (wave fms): at step k the owner of row k publishes x_k in
// LDS, and after one barrier every wave applies it to its rows below k.
constexpr int SWEEP_FWD_WAVES = SWEEP_BLOCK / 2;
template <int W>
__global__ __launch_bounds__(64 * SWEEP_FWD_WAVES) void sweep_fwd_diag_wave_kernel(const uint32_t *L, int64_t n,
                                                                                   uint32_t *X, int64_t ldx,
                                                                                   int64_t k0, int64_t k1) {
    extern __shared__ uint32_t tri_sh[];
    __shared__ uint32_t xk_sh[2][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.x * ldx * (W + 2);
    const int b = (int)(k1 - k0), wave = threadIdx.x >> 6;
    stage_block_tri<W>(L, n, k0, b, tri_sh);
    const int ra = 2 * wave, rb = 2 * wave + 1;  // this wave's rows of the block
    wmp::WV<W> va = wmp::zero<W>(), vb = wmp::zero<W>();
    if (ra < b) va = wmp::load<W>(rec_at<W>((const uint32_t *)x, k0 + ra));
    if (rb < b) vb = wmp::load<W>(rec_at<W>((const uint32_t *)x, k0 + rb));
    __syncthreads();
    for (int k = 0; k < b; k++) {
        const int par = k & 1;
        if (k == ra) wmp::store(xk_sh[par], va);  // final: every k' < k applied
        if (k == rb) wmp::store(xk_sh[par], vb);
        __syncthreads();
        if (rb > k && rb < b) {
            const wmp::WV<W> xk = wmp::load<W>(xk_sh[par]);
            if (ra > k) va = wmp::fms<W>(va, wmp::load<W>(tri_sh + (ra * (ra - 1) / 2 + k) * (W + 2)), xk);
            vb = wmp::fms<W>(vb, wmp::load<W>(tri_sh + (rb * (rb - 1) / 2 + k) * (W + 2)), xk);
        }
    }
    if (ra < b) wmp::store(rec_at<W>(x, k0 + ra), va);
    if (rb < b) wmp::store(rec_at<W>(x, k0 + rb), vb);
}

// v = RN(... RN(v - l_0 xb_0) ... - l_(b-1) xb_(b-1)), l_u at L + (base + u step)
// records, four fetches in flight
template <int W>
__device__ __forceinline__ void apply_block(Val<W> &v, const uint32_t *L, int64_t base, int64_t step,
                                            const uint32_t (*xb)[W + 2], int b) {
    Val<W> r0, r1, r2, r3, xk, t;
    auto fetch = [&](int u, Val<W> &dst) {
        if (u < b) load(rec_at<W>(L, base + u * step), dst);
    };
    auto apply = [&](int u, const Val<W> &l) {
        if (u < b) {
            load((const uint32_t *)xb[u], xk);
            fms(v, l, xk, t);
            v = t;
        }
    };
    fetch(0, r0);
    fetch(1, r1);
    fetch(2, r2);
    fetch(3, r3);
    for (int u = 0; u < b; u += 4) {
        apply(u, r0);
        fetch(u + 4, r0);
        apply(u + 1, r1);
        fetch(u + 5, r1);
        apply(u + 2, r2);
        fetch(u + 6, r2);
        apply(u + 3, r3);
        fetch(u + 7, r3);
    }
}

template <int W>
__global__ __launch_bounds__(64) void sweep_fwd_update_kernel(const uint32_t *L, int64_t n, uint32_t *X,
                                                              int64_t ldx, int64_t k0, int64_t k1) {
    __shared__ uint32_t xb[SWEEP_BLOCK][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.y * ldx * (W + 2);
    const int b = (int)(k1 - k0);
    for (int u = threadIdx.x; u < b * (W + 2); u += 64) xb[u / (W + 2)][u % (W + 2)] = rec_at<W>(x, k0)[u];
    __syncthreads();
    const int64_t i = k1 + (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    Val<W> v;
    load(rec_at<W>((const uint32_t *)x, i), v);
    apply_block<W>(v, L, i + k0 * n, n, xb, b);  // L(i, k0 + u)
    store(rec_at<W>(x, i), v);
}

template <int W>
__global__ __launch_bounds__(64) void sweep_bwd_diag_kernel(const uint32_t *L, int64_t n, uint32_t *X, int64_t ldx,
                                                            int64_t r0, int64_t r1) {
    extern __shared__ uint32_t tri_sh[];
    __shared__ uint32_t xb[SWEEP_BLOCK][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.x * ldx * (W + 2);
    const int b = (int)(r1 - r0), lane = threadIdx.x;
    stage_block_tri<W>(L, n, r0, b, tri_sh);
    Val<W> v;
    if (lane < b) load(rec_at<W>((const uint32_t *)x, r0 + lane), v);
    __syncthreads();
    for (int row = b - 1; row >= 0; row--) {
        if (lane == row) {
            Val<W> l, xk, t;
            for (int k = row + 1; k < b; k++) {  // L(k, row): entry (k, row) of the block's triangle
                load((const uint32_t *)(tri_sh + (k * (k - 1) / 2 + row) * (W + 2)), l);
                load((const uint32_t *)xb[k], xk);
                fms(v, l, xk, t);
                v = t;
            }
            store(xb[row], v);
        }
        __syncthreads();
    }
    if (lane < b) store(rec_at<W>(x, r0 + lane), v);
}

// The same on the whole wave: each single-rounding update of the block's
// rows is one wave fms (wmp::fms, the register fms's bits), the block's
// triangle and finished rows in LDS, operands fetched a step ahead.
template <int W>
__global__ __launch_bounds__(64) void sweep_bwd_diag_wave_kernel(const uint32_t *L, int64_t n, uint32_t *X,
                                                                 int64_t ldx, int64_t r0, int64_t r1) {
    extern __shared__ uint32_t tri_sh[];
    __shared__ uint32_t xb[SWEEP_BLOCK][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.x * ldx * (W + 2);
    const int b = (int)(r1 - r0);
    stage_block_tri<W>(L, n, r0, b, tri_sh);
    __syncthreads();
    for (int row = b - 1; row >= 0; row--) {
        wmp::WV<W> v = wmp::load<W>(rec_at<W>((const uint32_t *)x, r0 + row));
        uint32_t lw = 0u, xw = 0u;
        if (row + 1 < b) {
            lw = wmp::load_word<W>(tri_sh + ((row + 1) * row / 2 + row) * (W + 2));
            xw = wmp::load_word<W>(xb[row + 1]);
        }
        for (int k = row + 1; k < b; k++) {  // L(k, row) at entry (k, row) of the triangle, k ascending
            const wmp::WV<W> l = wmp::from_word<W>(lw), xk = wmp::from_word<W>(xw);
            if (k + 1 < b) {
                lw = wmp::load_word<W>(tri_sh + ((k + 1) * k / 2 + row) * (W + 2));
                xw = wmp::load_word<W>(xb[k + 1]);
            }
            v = wmp::fms<W>(v, l, xk);
        }
        wmp::store(xb[row], v);
        __threadfence_block();  // the next rows read this record's words from other lanes
        wmp::store(rec_at<W>(x, r0 + row), v);
    }
}

// Update of one row per wave (wave fms): row i applies the block's b
// finished values ascending, L's entry of each step fetched a step ahead.
// Forward: rows i >= k1 take L(i, k0 + u); backward: rows i < k0 take
// L(k0 + u, i).
template <int W>
__global__ __launch_bounds__(64) void sweep_update_wave_kernel(const uint32_t *L, int64_t n, uint32_t *X, int64_t ldx,
                                                               int64_t k0, int64_t k1, int backward) {
    __shared__ uint32_t xb[SWEEP_BLOCK][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.y * ldx * (W + 2);
    const int b = (int)(k1 - k0);
    for (int u = threadIdx.x; u < b * (W + 2); u += 64) xb[u / (W + 2)][u % (W + 2)] = rec_at<W>(x, k0)[u];
    __syncthreads();
    const int64_t i = backward ? (int64_t)blockIdx.x : k1 + (int64_t)blockIdx.x;
    const int64_t base = backward ? k0 + i * n : i + k0 * n, step = backward ? 1 : n;
    wmp::WV<W> v = wmp::load<W>(rec_at<W>((const uint32_t *)x, i));
    uint32_t lw = wmp::load_word<W>(rec_at<W>(L, base));
    for (int u = 0; u < b; u++) {
        const wmp::WV<W> l = wmp::from_word<W>(lw);
        if (u + 1 < b) lw = wmp::load_word<W>(rec_at<W>(L, base + (u + 1) * step));
        v = wmp::fms<W>(v, l, wmp::load<W>(xb[u]));
    }
    wmp::store(rec_at<W>(x, i), v);
}

template <int W>
__global__ __launch_bounds__(64) void sweep_bwd_update_kernel(const uint32_t *L, int64_t n, uint32_t *X,
                                                              int64_t ldx, int64_t r0, int64_t r1) {
    __shared__ uint32_t xb[SWEEP_BLOCK][W + 2];
    uint32_t *x = X + (int64_t)blockIdx.y * ldx * (W + 2);
    const int b = (int)(r1 - r0);
    for (int u = threadIdx.x; u < b * (W + 2); u += 64) xb[u / (W + 2)][u % (W + 2)] = rec_at<W>(x, r0)[u];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= r0) return;
    Val<W> v;
    load(rec_at<W>((const uint32_t *)x, i), v);
    apply_block<W>(v, L, r0 + i * n, 1, xb, b);  // L(r0 + u, i)
    store(rec_at<W>(x, i), v);
}

// ---------------------------------------------------------------------------
// Helpers of the exact multilinear forms (exact_forms.cpp)
// ---------------------------------------------------------------------------
// The exact product d x as hi + lo: hi = RN(d x), lo = x d - hi = -RN(hi -
// d x) (exactly representable: the error of a rounded product), for an
// m x n block of X (ldx) with d per row (mode 0), per column (1) or one
// value (2); hi and lo m x n with ld ldo.
template <int W>
__global__ void two_product_kernel(const uint32_t *X, int64_t ldx, const uint32_t *D, int mode, uint32_t *Hi,
                                   uint32_t *Lo, int64_t ldo, int64_t m, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    Val<W> x, d, hi, r;
    load(rec_at<W>(X, i + j * ldx), x);
    load(rec_at<W>(D, mode == 0 ? i : (mode == 1 ? j : 0)), d);
    mul(d, x, hi);
    fms(hi, d, x, r);  // RN(hi - d x), exact
    if (r.kind == mpd::KIND_REG || r.kind == -mpd::KIND_REG) r.kind = -r.kind;
    store(rec_at<W>(Hi, i + j * ldo), hi);
    store(rec_at<W>(Lo, i + j * ldo), r);
}

// out (h x h, ld h) = the symmetric matrix of a column-packed upper triangle
// (entry (i, j), i <= j, at j (j + 1) / 2 + i), times `scale` off the
// diagonal when scale is given (an exact power of two or one value)
template <int W>
__global__ void unpack_symmetric_kernel(const uint32_t *P, int64_t h, uint32_t *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= h * h) return;
    const int64_t i = e % h, j = e / h;
    const int64_t a = i < j ? i : j, b = i < j ? j : i;
    Val<W> v;
    load(rec_at<W>(P, b * (b + 1) / 2 + a), v);
    store(rec_at<W>(out, e), v);
}

// The upper-triangular weight matrix of svec_quadratic as two exact parts:
// from the svec x (entry (i, j), i <= j, at j (j + 1) / 2 + i), M_ii = x_ii
// (lo 0) and, for i < j, s x_ij = hi + lo exactly; zero below the diagonal
template <int W>
__global__ void quadratic_weights_kernel(const uint32_t *P, int64_t h, const uint32_t *S, uint32_t *Hi, uint32_t *Lo) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= h * h) return;
    const int64_t i = e % h, j = e / h;
    Val<W> hi, lo;
    set_zero(hi);
    set_zero(lo);
    if (i == j) {
        load(rec_at<W>(P, j * (j + 1) / 2 + i), hi);
    } else if (i < j) {
        Val<W> x, s, r;
        load(rec_at<W>(P, j * (j + 1) / 2 + i), x);
        load(S, s);
        mul(s, x, hi);
        fms(hi, s, x, r);
        if (r.kind == mpd::KIND_REG || r.kind == -mpd::KIND_REG) r.kind = -r.kind;
        lo = r;
    }
    store(rec_at<W>(Hi, e), hi);
    store(rec_at<W>(Lo, e), lo);
}

// |x| in place (count values)
template <int W>
__global__ void abs_values_kernel(uint32_t *X, int64_t count) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    uint32_t *k = rec_at<W>(X, e) + W;
    const int32_t kind = (int32_t)*k;
    if (kind < 0) *k = (uint32_t)(-kind);
}

// c = axpby(alpha, x, beta, c) entrywise, the reference's MPFR axpby
// (blas/mpfr.rs:61-80): p = 0, x, -x or RN(alpha x) for alpha codes 0 (zero),
// 1 (one), 2 (minus one), 3 (other); then p, RN(p + c), RN(p - c) or
// RN(p + RN(beta c)) for beta's codes.  x = P[i + j ldp] (read from
// P[min + max ldp] when sym: a symmetric product held as its upper
// triangle); tri 0 every entry, 1 i <= j, 2 i >= j; packed: c holds the
// column-packed upper triangle.
template <int W>
__global__ void axpby_kernel(const uint32_t *P, int64_t ldp, int64_t m, int64_t n, int acode, const uint32_t *alpha,
                             int bcode, const uint32_t *beta, uint32_t *C, int64_t ldc, int tri, int packed,
                             int sym) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    if ((tri == 1 && i > j) || (tri == 2 && i < j)) return;
    uint32_t *c = rec_at<W>(C, packed ? j * (j + 1) / 2 + i : i + j * ldc);
    Val<W> p, y, t;
    set_zero(p);
    if (acode != 0) {
        const int64_t pi = sym && i > j ? j : i, pj = sym && i > j ? i : j;
        load(rec_at<W>(P, pi + pj * ldp), p);
        if (acode == 2) {
            if (is_regular(p) || p.kind == mpd::KIND_INF || p.kind == -mpd::KIND_INF) p.kind = -p.kind;
        } else if (acode == 3) {
            Val<W> a;
            load(alpha, a);
            mul(a, p, t);
            p = t;
        }
    }
    if (bcode != 0) {
        load(c, y);
        if (bcode == 3) {
            Val<W> b;
            load(beta, b);
            mul(b, y, t);
            y = t;
        }
        add_signed(p, y, bcode == 2, t);
        p = t;
    }
    store(c, p);
}

// Grouped exact blocks on the device: block b's y (rows x ncols, ld rows)
// into Y and Z = RN(d_r y) into Z, rows row0.. of the dense K x m operands
// (ld K) at its global columns; one workgroup row per block
template <int W>
__global__ void block_scatter_kernel(const BlockScatter *descs, uint32_t *Z, uint32_t *Y, int64_t K) {
    const BlockScatter b = descs[blockIdx.y];
    const int64_t total = b.rows * b.ncols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % b.rows, c = e / b.rows;
        const int64_t dst = (b.row0 + r) + b.cols[c] * K;
        Val<W> y, d, z;
        load(rec_at<W>((const uint32_t *)b.y, e), y);
        load(rec_at<W>((const uint32_t *)b.d, r), d);
        mul(d, y, z);
        store(rec_at<W>(Y, dst), y);
        store(rec_at<W>(Z, dst), z);
    }
}

// out[p] = G[ia[p] + ib[p] ldg]
template <int W>
__global__ void gather_kernel(const uint32_t *G, int64_t ldg, const int64_t *ia, const int64_t *ib, int64_t count,
                              uint32_t *out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    Val<W> v;
    load(rec_at<W>(G, ia[p] + ib[p] * ldg), v);
    store(rec_at<W>(out, p), v);
}

// count zero values
template <int W>
__global__ void zero_values_kernel(uint32_t *P, int64_t count) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    Val<W> z;
    set_zero(z);
    store(rec_at<W>(P, e), z);
}

// X[i + j ldx] = RN(RN(Wm[i + j ldw] - S[i + j lds]) D[i]) (second_many's
// update, arrow.rs:1717-1729) for an m x n block
template <int W>
__global__ void arrow_second_kernel(const uint32_t *Wm, int64_t ldw, const uint32_t *S, int64_t lds,
                                    const uint32_t *D, uint32_t *X, int64_t ldx, int64_t m, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    Val<W> w, sv, t, di, r;
    load(rec_at<W>(Wm, i + j * ldw), w);
    load(rec_at<W>(S, i + j * lds), sv);
    add_signed(w, sv, true, t);
    load(rec_at<W>(D, i), di);
    mul(t, di, r);
    store(rec_at<W>(X, i + j * ldx), r);
}

// S[i + j*lds] = RN(S - P[i + j*ldp]) for an m x n block
template <int W>
__global__ void ew_sub_rect_kernel(uint32_t *S, int64_t lds, const uint32_t *P, int64_t ldp,
                                   int64_t m, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t i = e % m, j = e / m;
    Val<W> a, b, r;
    load(rec_at<W>((const uint32_t *)S, i + j * lds), a);
    load(rec_at<W>(P, i + j * ldp), b);
    add_signed(a, b, true, r);
    store(rec_at<W>(S, i + j * lds), r);
}

// first nonzero column of each row of the lower triangle
template <int W>
__global__ void envelope_kernel(const uint32_t *A, int64_t ld, int64_t n, int32_t *first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t k = 0;
    while (k < i) {
        const int32_t kind = (int32_t)rec_at<W>(A, i + k * ld)[W];
        if (kind != mpd::KIND_ZERO && kind != -mpd::KIND_ZERO) break;
        k++;
    }
    first[i] = (int32_t)k;
}

// mirror the upper triangle onto the lower one (values)
template <int W>
__global__ void mirror_upper_kernel(uint32_t *S, int64_t n, int64_t ld) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    if (i <= j) return;
    const uint32_t *src = rec_at<W>((const uint32_t *)S, j + i * ld);
    uint32_t *dst = rec_at<W>(S, i + j * ld);
    for (int u = 0; u < W + 2; u++) dst[u] = src[u];
}

template <int W>
__global__ void mirror_lower_kernel(uint32_t *S, int64_t n, int64_t ld) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e % n, j = e / n;
    if (i <= j) return;
    const uint32_t *src = rec_at<W>((const uint32_t *)S, i + j * ld);
    uint32_t *dst = rec_at<W>(S, j + i * ld);
    for (int u = 0; u < W + 2; u++) dst[u] = src[u];
}

// |a| > |b| for finite records (zeros of either sign smallest)
template <int W>
__device__ __forceinline__ bool abs_greater(const uint32_t *a, const uint32_t *b) {
    const int32_t ka = (int32_t)a[W], kb = (int32_t)b[W];
    if (ka == mpd::KIND_ZERO || ka == -mpd::KIND_ZERO) return false;
    if (kb == mpd::KIND_ZERO || kb == -mpd::KIND_ZERO) return true;
    const int32_t ea = (int32_t)a[W + 1], eb = (int32_t)b[W + 1];
    if (ea != eb) return ea > eb;
    for (int i = W - 1; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return false;
}

// out[c] = max_i |X[i + c ldx]| over m rows (the infinity norm of column c),
// +infinity when the column holds a NaN or an infinity; one workgroup of
// 256 threads per column, the largest entry's index reduced in LDS.
template <int W>
__global__ void __launch_bounds__(256) col_absmax_kernel(const uint32_t *X, int64_t ldx, int64_t m, uint32_t *out) {
    __shared__ int64_t best[256];
    __shared__ int nonfinite;
    const uint32_t *col = rec_at<W>(X, (int64_t)blockIdx.x * ldx);
    if (threadIdx.x == 0) nonfinite = 0;
    __syncthreads();
    int64_t b = -1;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t *r = rec_at<W>(col, i);
        const int32_t k = (int32_t)r[W];
        if (k != mpd::KIND_ZERO && k != -mpd::KIND_ZERO && k != mpd::KIND_REG && k != -mpd::KIND_REG) {
            nonfinite = 1;
            continue;
        }
        if (b < 0 || abs_greater<W>(r, rec_at<W>(col, b))) b = i;
    }
    best[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int64_t o = best[threadIdx.x + s], cur = best[threadIdx.x];
            if (o >= 0 && (cur < 0 || abs_greater<W>(rec_at<W>(col, o), rec_at<W>(col, cur)))) best[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    uint32_t *dst = rec_at<W>(out, (int64_t)blockIdx.x);
    const int64_t i = best[0];
    if (nonfinite || i < 0) {
        for (int u = 0; u < W; u++) dst[u] = 0;
        dst[W] = (uint32_t)(nonfinite ? mpd::KIND_INF : mpd::KIND_ZERO);
        dst[W + 1] = 0;
        return;
    }
    const uint32_t *src = rec_at<W>(col, i);
    for (int u = 0; u < W + 2; u++) dst[u] = src[u];
    const int32_t k = (int32_t)src[W];
    dst[W] = (uint32_t)(k < 0 ? -k : k);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
template <int W>
size_t slot_bytes(int threads) {
    return (size_t)Win<W>::SLOTS * threads * sizeof(int64_t);
}
// dynamic LDS of a panel kernel (its slot windows), raising the kernel's
// limit above the 64 KiB default where needed
template <int W>
size_t panel_lds(const void *kernel, int threads) {
    const size_t bytes = slot_bytes<W>(threads);
    if (bytes > 65536)
        HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return bytes;
}
inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
}  // namespace

#define SDPX_DISPATCH_W(limbs, ...)                                     \
    switch (limbs) {                                                    \
        case 2: { constexpr int W = 4; __VA_ARGS__; break; }            \
        case 4: { constexpr int W = 8; __VA_ARGS__; break; }            \
        case 8: { constexpr int W = 16; __VA_ARGS__; break; }           \
        default: throw Error(SDPX_GPU_INVALID, "GPU panels support 128, 256 and 512 bits"); \
    }

// Lanes per exact dot in the latency-bound panel kernels (the rows of one
// factor column, the rows of a triangular solve): 64 (one wave a dot) unless
// SDPX_GPU_DOT_LANES selects 16 or 32.  The forward solve over many leaf
// columns keeps 16-lane groups (four dots a wave).
namespace {
int dot_lanes() {
    static const int g = [] {
        const char *e = std::getenv("SDPX_GPU_DOT_LANES");
        const int v = e ? std::atoi(e) : 64;
        return v == 16 || v == 32 ? v : 64;
    }();
    return g;
}
}  // namespace
#define SDPX_DISPATCH_G(...)                                \
    switch (dot_lanes()) {                                  \
        case 16: { constexpr int G = 16; __VA_ARGS__; break; } \
        case 32: { constexpr int G = 32; __VA_ARGS__; break; } \
        default: { constexpr int G = 64; __VA_ARGS__; break; } \
    }

bool mp_kernels_support(int limbs) { return limbs == 2 || limbs == 4 || limbs == 8; }

void mp_ew_sub_packed(int limbs, void *S, int64_t ld, const void *P, int64_t n, int lower,
                      hipStream_t st) {
    const int64_t total = n * (n + 1) / 2;
    if (total == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(ew_sub_packed_kernel<W>, dim3(blocks_for(total, 256)),
                                              dim3(256), 0, st, (uint32_t *)S, ld,
                                              (const uint32_t *)P, n, lower));
    HIP_CHECK(hipGetLastError());
}

void mp_ew_mul_diag(int limbs, const void *Y, int64_t ldy, const void *D, void *Z, int64_t ldz,
                    int64_t m, int64_t n, bool per_row, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(ew_mul_diag_kernel<W>, dim3(blocks_for(m * n, 256)),
                                              dim3(256), 0, st, (const uint32_t *)Y, ldy,
                                              (const uint32_t *)D, (uint32_t *)Z, ldz, m, n,
                                              per_row ? 1 : 0));
    HIP_CHECK(hipGetLastError());
}

void mp_sampled_entries(int limbs, const void *G, const void *w, int64_t K, void *H,
                        hipStream_t st) {
    if (K == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(sampled_entries_kernel<W>, dim3(blocks_for(K * K, 256)),
                                              dim3(256), 0, st, (const uint32_t *)G,
                                              (const uint32_t *)w, K, (uint32_t *)H));
    HIP_CHECK(hipGetLastError());
}

void mp_exp_ranges(int limbs, const void *L, int64_t ldl, int64_t n, int mode, void *out,
                   hipStream_t st) {
    if (n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(exp_ranges_kernel<W>, dim3((unsigned)n), dim3(64), 0, st,
                                              (const uint32_t *)L, ldl, n, mode, (int2 *)out));
    HIP_CHECK(hipGetLastError());
}

// tuning switch "fwd_group": lanes per exact dot of the forward rows (4, 8
// (default), 16 or 32)
static std::atomic<int> g_fwd_group{8};

void mp_forward_rows(int limbs, const void *L_all, const int32_t *first_all, void *Y_all,
                     const void *ranges_all, const FwdLeaf *leaves, int nleaves, int64_t t, int *fail,
                     hipStream_t st) {
    if (nleaves == 0 || t == 0) return;
    const int g = g_fwd_group.load();
    SDPX_DISPATCH_W(limbs, {
        auto kern = g == 4 ? forward_rows_kernel<W, 4>
                           : (g == 8 ? forward_rows_kernel<W, 8>
                                     : (g == 32 ? forward_rows_kernel<W, 32> : forward_rows_kernel<W, 16>));
        const int64_t G = g;
        const size_t lds = panel_lds<W>((const void *)kern, ROW_THREADS);
        hipLaunchKernelGGL(kern, dim3(blocks_for(t * G, ROW_THREADS), nleaves), dim3(ROW_THREADS), lds, st,
                           (const uint32_t *)L_all, first_all, (uint32_t *)Y_all, (const int2 *)ranges_all,
                           leaves, t, fail);
    });
    HIP_CHECK(hipGetLastError());
}

// Workgroups of `lds` bytes that the device holds at once (CUs x per-CU LDS).
static int64_t rows_resident(size_t lds) {
    static int64_t cus = 0, lds_cu = 0;
    if (!cus) {
        int dev = 0, v = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        cus = v;
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        lds_cu = v > 0 ? v : 65536;
    }
    return cus * std::max<int64_t>(1, lds_cu / (int64_t)lds);
}

// Tuning switches (sdpx_gpu_set_tuning): "trsm_rl" 0/1 turns the
// right-looking pass of the triangular-solve base off/on (default on, or
// SDPX_GPU_TRSM_RL=0 in the environment), "trsm_rl_room" sets its window's
// room below the first term in bits (default 128; a negative room sends
// every row to the left-looking kernel).
static std::atomic<int> g_trsm_rl{-1};
static std::atomic<int64_t> g_trsm_room{128};
static bool trsm_right_looking() {
    const int v = g_trsm_rl.load();
    if (v >= 0) return v != 0;
    static const bool env = [] {
        const char *e = std::getenv("SDPX_GPU_TRSM_RL");
        return !(e && e[0] == '0');
    }();
    return env;
}
// "ldl_one_launch": 1 (default) a single ldl_dots block (the recursive
// factor's bases) in one launch of ldl_rows_kernel (rows ordered by flags),
// several blocks (the leaves) one launch per column; 2 every call in one
// launch of row workgroups (all blocks' rows interleaved); 0 column launches
// only.
static std::atomic<int> g_ldl_one_launch{1};
// "sweep_wave" 1 (default): the border's diagonal blocks (forward and
// backward) and the blocks' updates of the other rows on whole waves (wave
// fms); 0: one lane a row.
static std::atomic<int> g_sweep_wave{1};
static int ldl_one_launch() { return g_ldl_one_launch.load(); }
// rows of one launch of ldl_rows_kernel at most (larger blocks take the
// column launches)
constexpr int64_t LDL_ROWS_MAX = 4096;

bool mp_set_tuning(const char *name, int64_t value) {
    if (!std::strcmp(name, "ldl_one_launch")) {
        if (value < 0 || value > 2) return false;
        g_ldl_one_launch = (int)value;
        return true;
    }
    if (!std::strcmp(name, "fwd_group")) {
        if (value != 4 && value != 8 && value != 16 && value != 32) return false;
        g_fwd_group = (int)value;
        return true;
    }
    if (!std::strcmp(name, "sweep_wave")) {
        g_sweep_wave = value ? 1 : 0;
        return true;
    }
    if (!std::strcmp(name, "trsm_rl")) {
        g_trsm_rl = value ? 1 : 0;
        return true;
    }
    if (!std::strcmp(name, "trsm_rl_room")) {
        g_trsm_room = value;
        return true;
    }
    return false;
}

void mp_trsm_base(int limbs, const void *L, int64_t ldl, int64_t n, void *B, int64_t ldb,
                  int64_t m, const int32_t *low, const void *lranges, int *fail, void *work,
                  hipStream_t st) {
    if (m == 0 || n == 0) return;
    int *fb_count = (int *)work;
    int32_t *fb_rows = (int32_t *)work + 1;
    SDPX_DISPATCH_W(limbs, {
        // the right-looking form holds a window per column (twice the LDS of
        // a left-looking row): it runs when all m rows fit on the chip at once
        const size_t lds = slot_bytes<W>(RL_COLS) + RL_COLS * sizeof(int64_t);
        const bool rl = trsm_right_looking() && n <= 128 && m <= rows_resident(lds);
        if (rl) {
            HIP_CHECK(hipMemsetAsync(fb_count, 0, sizeof(int), st));
            auto kr = trsm_rl_kernel<W>;
            if (lds > 65536)
                HIP_CHECK(hipFuncSetAttribute((const void *)kr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(kr, dim3((unsigned)m), dim3(RL_THREADS), lds, st, (const uint32_t *)L, ldl, n,
                               (uint32_t *)B, ldb, m, low, (const int2 *)lranges, (int64_t)g_trsm_room.load(),
                               fb_count, fb_rows);
        }
        SDPX_DISPATCH_G({
            auto kern = G == 64 && n <= 128 ? trsm_base_wave_kernel<W> : trsm_base_kernel<W, G>;
            const size_t lds = panel_lds<W>((const void *)kern, ROW_THREADS);
            hipLaunchKernelGGL(kern, dim3(blocks_for(m * G, ROW_THREADS)), dim3(ROW_THREADS), lds, st,
                               (const uint32_t *)L, ldl, n, (uint32_t *)B, ldb, m, low, (const int2 *)lranges,
                               fail, rl ? fb_rows : nullptr, rl ? fb_count : nullptr);
        });
    });
    HIP_CHECK(hipGetLastError());
}

// SDPX_GPU_LDL_ONE_WORKGROUP=1: every block's columns on one workgroup
// (the single-launch form), for comparison.
static bool ldl_one_workgroup() {
    static const bool v = [] {
        const char *e = std::getenv("SDPX_GPU_LDL_ONE_WORKGROUP");
        return e && e[0] == '1';
    }();
    return v;
}

void mp_ldl_dots(int limbs, void *A, void *dinv, const int32_t *first, const int8_t *signs,
                 int sign_mode, const void *blocks, int nblocks, int64_t nmax, void *scratch,
                 const void *reg_eps, const void *reg_delta, int *status, hipStream_t st) {
    if (nblocks == 0 || nmax == 0) return;
    if (ldl_one_workgroup() && nmax <= LDL_STATS_MAX) {
        SDPX_DISPATCH_W(limbs, {
            auto kern = ldl_dots_kernel<W, DOT_G>;
            constexpr int threads = ldl_threads<W>();
            const size_t lds = slot_bytes<W>(threads) + 4 * LDL_STATS_MAX * sizeof(int32_t);
            if (lds > 65536)
                HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(kern, dim3(nblocks), dim3(threads), lds, st, (uint32_t *)A,
                               (uint32_t *)dinv, first, signs, sign_mode, (const LdlBlock *)blocks,
                               (uint32_t *)scratch, (const uint32_t *)reg_eps,
                               (const uint32_t *)reg_delta, status);
        });
        HIP_CHECK(hipGetLastError());
        return;
    }
    // one launch of row workgroups (rows of all blocks interleaved) when
    // ldl_one_launch is on (1: one block, the recursive factor's base blocks;
    // 2: also several blocks, the leaves); else one launch per column
    if ((ldl_one_launch() == 2 || (ldl_one_launch() && nblocks == 1)) && nmax <= LDL_ROWS_MAX) {
        SDPX_DISPATCH_W(limbs, {
            hipLaunchKernelGGL(ldl_flags_clear_kernel<W>, dim3(blocks_for(2 * nmax, 256), nblocks), dim3(256), 0, st,
                               (uint32_t *)A, (uint32_t *)dinv, first, (uint32_t *)scratch, (const LdlBlock *)blocks);
            auto kern = ldl_rows_kernel<W>;
            const size_t lds = panel_lds<W>((const void *)kern, 2 * ROW_THREADS);  // row windows, look-ahead windows
            hipLaunchKernelGGL(kern, dim3((unsigned)(nmax * nblocks)), dim3(3 * ROW_THREADS), lds, st, (uint32_t *)A,
                               (uint32_t *)dinv, first, signs, sign_mode, (const LdlBlock *)blocks, nblocks,
                               (uint32_t *)scratch, (const uint32_t *)reg_eps, (const uint32_t *)reg_delta, status);
            hipLaunchKernelGGL(ldl_col_writeout_kernel<W>, dim3(blocks_for(nmax * nmax, 256), nblocks), dim3(256),
                               0, st, (uint32_t *)A, (uint32_t *)dinv, (const LdlBlock *)blocks, (uint32_t *)scratch,
                               (const int *)status);
        });
        HIP_CHECK(hipGetLastError());
        return;
    }
    SDPX_DISPATCH_W(limbs, {
        SDPX_DISPATCH_G({
            auto kern = G == 64 ? ldl_col_wave_kernel<W> : ldl_col_kernel<W, G>;
            const size_t lds = panel_lds<W>((const void *)kern, 2 * ROW_THREADS);  // row groups, pivot windows
            for (int64_t j = -1; j < nmax - 1; j++) {
                const int64_t rows = j < 0 ? 1 : nmax - j - 1;
                hipLaunchKernelGGL(kern, dim3(blocks_for(rows * G, ROW_THREADS), nblocks), dim3(3 * ROW_THREADS), lds,
                                   st, (uint32_t *)A, (uint32_t *)dinv, first, signs, sign_mode,
                                   (const LdlBlock *)blocks, (uint32_t *)scratch, (const uint32_t *)reg_eps,
                                   (const uint32_t *)reg_delta, status, j);
            }
        });
        hipLaunchKernelGGL(ldl_col_writeout_kernel<W>, dim3(blocks_for(nmax * nmax, 256), nblocks), dim3(256),
                           0, st, (uint32_t *)A, (uint32_t *)dinv, (const LdlBlock *)blocks, (uint32_t *)scratch,
                           (const int *)status);
    });
    HIP_CHECK(hipGetLastError());
}

void mp_ew_sub_rect(int limbs, void *S, int64_t lds, const void *P, int64_t ldp, int64_t m,
                    int64_t n, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(ew_sub_rect_kernel<W>, dim3(blocks_for(m * n, 256)),
                                              dim3(256), 0, st, (uint32_t *)S, lds,
                                              (const uint32_t *)P, ldp, m, n));
    HIP_CHECK(hipGetLastError());
}

namespace {
// rows per thread of the sweep kernels for an order-n factor, and threads
int sweep_rows(int64_t n, int &threads) {
    threads = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (n + 63) / 64 * 64));
    const int64_t r = (n + threads - 1) / threads;
    return r <= 1 ? 1 : (r <= 2 ? 2 : (r <= 4 ? 4 : (r <= 8 ? 8 : 0)));
}
}  // namespace

#define SDPX_DISPATCH_R(R_, ...)                                    \
    switch (R_) {                                                  \
        case 1: { constexpr int R = 1; __VA_ARGS__; break; }       \
        case 2: { constexpr int R = 2; __VA_ARGS__; break; }       \
        case 4: { constexpr int R = 4; __VA_ARGS__; break; }       \
        default: { constexpr int R = 8; __VA_ARGS__; break; }      \
    }

bool mp_sweep_supports(int64_t n) {
    int threads;
    return sweep_rows(n, threads) != 0;
}

void mp_arrow_forward(int limbs, const void *L_all, const void *dinv_all, const void *descs, int ndesc,
                      int64_t nmax, int64_t cols, void *X, void *Wout, void *Vout, int64_t ldw, hipStream_t st) {
    if (ndesc == 0 || nmax == 0 || cols == 0) return;
    int threads;
    const int R = sweep_rows(nmax, threads);
    if (R == 0) throw Error(SDPX_GPU_INVALID, "sweep order above 8192");
    SDPX_DISPATCH_W(limbs, SDPX_DISPATCH_R(R, hipLaunchKernelGGL((arrow_forward_kernel<W, R>),
        dim3((unsigned)cols, ndesc), dim3(threads), 0, st, (const uint32_t *)L_all, (const uint32_t *)dinv_all,
        (const SweepDesc *)descs, (uint32_t *)X, (uint32_t *)Wout, (uint32_t *)Vout, ldw)));
    HIP_CHECK(hipGetLastError());
}

void mp_arrow_backward(int limbs, const void *L_all, const void *descs, int ndesc, int64_t nmax, int64_t cols,
                       void *X, int back, hipStream_t st) {
    if (ndesc == 0 || nmax == 0 || cols == 0) return;
    int threads;
    const int R = sweep_rows(nmax, threads);
    if (R == 0) throw Error(SDPX_GPU_INVALID, "sweep order above 8192");
    const size_t lds = (size_t)back * (value_words(limbs)) * 4;
    // the block's own rows on wave 0 in wave form when the workgroup has a
    // whole wave (threads >= 64)
    const int wave_chain = g_sweep_wave.load() && threads >= 64 ? 1 : 0;
    SDPX_DISPATCH_W(limbs, SDPX_DISPATCH_R(R, hipLaunchKernelGGL((arrow_backward_kernel<W, R>),
        dim3((unsigned)cols, ndesc), dim3(threads), lds, st, (const uint32_t *)L_all, (const SweepDesc *)descs,
        (uint32_t *)X, back, wave_chain)));
    HIP_CHECK(hipGetLastError());
}

void mp_sweep_blocked(int limbs, const void *L, int64_t n, int64_t cols, void *X, int64_t ldx, bool backward,
                      int back, hipStream_t st) {
    if (n == 0 || cols == 0) return;
    if (back < 1 || back > SWEEP_BLOCK) throw Error(SDPX_GPU_INVALID, "sweep block above 32 rows");
    SDPX_DISPATCH_W(limbs, {
        const size_t tri = (size_t)SWEEP_BLOCK * (SWEEP_BLOCK - 1) / 2 * (W + 2) * 4;
        if (!backward) {
            for (int64_t k0 = 0; k0 < n; k0 += SWEEP_BLOCK) {
                const int64_t k1 = std::min<int64_t>(n, k0 + SWEEP_BLOCK);
                if (g_sweep_wave.load())
                    hipLaunchKernelGGL(sweep_fwd_diag_wave_kernel<W>, dim3((unsigned)cols), dim3(64 * SWEEP_FWD_WAVES),
                                       tri, st, (const uint32_t *)L, n, (uint32_t *)X, ldx, k0, k1);
                else
                    hipLaunchKernelGGL(sweep_fwd_diag_kernel<W>, dim3((unsigned)cols), dim3(64), tri, st,
                                       (const uint32_t *)L, n, (uint32_t *)X, ldx, k0, k1);
                if (k1 < n && g_sweep_wave.load())
                    hipLaunchKernelGGL(sweep_update_wave_kernel<W>, dim3((unsigned)(n - k1), (unsigned)cols), dim3(64),
                                       0, st, (const uint32_t *)L, n, (uint32_t *)X, ldx, k0, k1, 0);
                else if (k1 < n)
                    hipLaunchKernelGGL(sweep_fwd_update_kernel<W>, dim3((unsigned)((n - k1 + 63) / 64), (unsigned)cols),
                                       dim3(64), 0, st, (const uint32_t *)L, n, (uint32_t *)X, ldx, k0, k1);
            }
        } else {
            for (int64_t r1 = n; r1 > 0;) {
                const int64_t r0 = r1 > back ? r1 - back : 0;
                if (g_sweep_wave.load())
                    hipLaunchKernelGGL(sweep_bwd_diag_wave_kernel<W>, dim3((unsigned)cols), dim3(64), tri, st,
                                       (const uint32_t *)L, n, (uint32_t *)X, ldx, r0, r1);
                else
                    hipLaunchKernelGGL(sweep_bwd_diag_kernel<W>, dim3((unsigned)cols), dim3(64), tri, st,
                                       (const uint32_t *)L, n, (uint32_t *)X, ldx, r0, r1);
                if (r0 > 0 && g_sweep_wave.load())
                    hipLaunchKernelGGL(sweep_update_wave_kernel<W>, dim3((unsigned)r0, (unsigned)cols), dim3(64), 0, st,
                                       (const uint32_t *)L, n, (uint32_t *)X, ldx, r0, r1, 1);
                else if (r0 > 0)
                    hipLaunchKernelGGL(sweep_bwd_update_kernel<W>, dim3((unsigned)((r0 + 63) / 64), (unsigned)cols),
                                       dim3(64), 0, st, (const uint32_t *)L, n, (uint32_t *)X, ldx, r0, r1);
                r1 = r0;
            }
        }
    });
    HIP_CHECK(hipGetLastError());
}

void mp_two_product(int limbs, const void *X, int64_t ldx, const void *D, int mode, void *Hi, void *Lo, int64_t ldo,
                    int64_t m, int64_t n, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(two_product_kernel<W>, dim3(blocks_for(m * n, 256)), dim3(256), 0, st,
                                              (const uint32_t *)X, ldx, (const uint32_t *)D, mode, (uint32_t *)Hi,
                                              (uint32_t *)Lo, ldo, m, n));
    HIP_CHECK(hipGetLastError());
}

void mp_quadratic_weights(int limbs, const void *P, int64_t h, const void *S, void *Hi, void *Lo, hipStream_t st) {
    if (h == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(quadratic_weights_kernel<W>, dim3(blocks_for(h * h, 256)), dim3(256), 0,
                                              st, (const uint32_t *)P, h, (const uint32_t *)S, (uint32_t *)Hi,
                                              (uint32_t *)Lo));
    HIP_CHECK(hipGetLastError());
}

void mp_unpack_symmetric(int limbs, const void *P, int64_t h, void *out, hipStream_t st) {
    if (h == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(unpack_symmetric_kernel<W>, dim3(blocks_for(h * h, 256)), dim3(256), 0,
                                              st, (const uint32_t *)P, h, (uint32_t *)out));
    HIP_CHECK(hipGetLastError());
}

void mp_abs_values(int limbs, void *X, int64_t count, hipStream_t st) {
    if (count == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(abs_values_kernel<W>, dim3(blocks_for(count, 256)), dim3(256), 0, st,
                                              (uint32_t *)X, count));
    HIP_CHECK(hipGetLastError());
}

void mp_axpby(int limbs, const void *P, int64_t ldp, int64_t m, int64_t n, int acode, const void *alpha, int bcode,
              const void *beta, void *C, int64_t ldc, int tri, bool packed, bool sym, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(axpby_kernel<W>, dim3(blocks_for(m * n, 256)), dim3(256), 0, st,
                                              (const uint32_t *)P, ldp, m, n, acode, (const uint32_t *)alpha, bcode,
                                              (const uint32_t *)beta, (uint32_t *)C, ldc, tri, packed ? 1 : 0,
                                              sym ? 1 : 0));
    HIP_CHECK(hipGetLastError());
}

void mp_block_scatter(int limbs, const BlockScatter *descs, int nblocks, int64_t max_elems, void *Z, void *Y,
                      int64_t K, hipStream_t st) {
    if (nblocks == 0 || max_elems == 0) return;
    const unsigned gx = (unsigned)std::min<int64_t>(blocks_for(max_elems, 256), 1024);
    for (int b0 = 0; b0 < nblocks; b0 += 65535) {
        const int nb = std::min(65535, nblocks - b0);
        SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(block_scatter_kernel<W>, dim3(gx, nb), dim3(256), 0, st, descs + b0,
                                                  (uint32_t *)Z, (uint32_t *)Y, K));
        HIP_CHECK(hipGetLastError());
    }
}

void mp_gather(int limbs, const void *G, int64_t ldg, const int64_t *ia, const int64_t *ib, int64_t count, void *out,
               hipStream_t st) {
    if (count == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(gather_kernel<W>, dim3(blocks_for(count, 256)), dim3(256), 0, st,
                                              (const uint32_t *)G, ldg, ia, ib, count, (uint32_t *)out));
    HIP_CHECK(hipGetLastError());
}

void mp_zero_values(int limbs, void *P, int64_t count, hipStream_t st) {
    if (count == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(zero_values_kernel<W>, dim3(blocks_for(count, 256)), dim3(256), 0, st,
                                              (uint32_t *)P, count));
    HIP_CHECK(hipGetLastError());
}

void mp_arrow_second(int limbs, const void *Wm, int64_t ldw, const void *S, int64_t lds, const void *D, void *X,
                     int64_t ldx, int64_t m, int64_t n, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(arrow_second_kernel<W>, dim3(blocks_for(m * n, 256)), dim3(256), 0,
                                              st, (const uint32_t *)Wm, ldw, (const uint32_t *)S, lds,
                                              (const uint32_t *)D, (uint32_t *)X, ldx, m, n));
    HIP_CHECK(hipGetLastError());
}

void mp_envelope(int limbs, const void *A, int64_t ld, int64_t n, int32_t *first, hipStream_t st) {
    if (n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(envelope_kernel<W>, dim3(blocks_for(n, 256)), dim3(256),
                                              0, st, (const uint32_t *)A, ld, n, first));
    HIP_CHECK(hipGetLastError());
}

void mp_trsm_base_n(int limbs, const void *L, int64_t ldl, int64_t n, void *B, int64_t ldb,
                    int64_t m, const void *lranges, int *fail, hipStream_t st) {
    if (m == 0 || n == 0) return;
    SDPX_DISPATCH_W(limbs, {
        SDPX_DISPATCH_G({
            auto kern = G == 64 && n <= 128 ? trsm_base_n_wave_kernel<W> : trsm_base_n_kernel<W, G>;
            const size_t lds = panel_lds<W>((const void *)kern, ROW_THREADS);
            hipLaunchKernelGGL(kern, dim3(blocks_for(m * G, ROW_THREADS)), dim3(ROW_THREADS), lds, st,
                               (const uint32_t *)L, ldl, n, (uint32_t *)B, ldb, m, (const int2 *)lranges, fail);
        });
    });
    HIP_CHECK(hipGetLastError());
}

void mp_transpose(int limbs, const void *in, int64_t ldi, int64_t m, int64_t n, void *out,
                  int64_t ldo, hipStream_t st) {
    if (m * n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(transpose_kernel<W>, dim3(blocks_for(m * n, 256)),
                                              dim3(256), 0, st, (const uint32_t *)in, ldi, m, n,
                                              (uint32_t *)out, ldo));
    HIP_CHECK(hipGetLastError());
}

void mp_potrf(int limbs, void *A, int64_t n, int64_t lda, int lower, int *status, hipStream_t st) {
    SDPX_DISPATCH_W(limbs, {
        SDPX_DISPATCH_G({
            auto kern = potrf_col_kernel<W, G>;
            const size_t lds = panel_lds<W>((const void *)kern, ROW_THREADS);
            for (int64_t j = -1; j < n - 1; j++)
                hipLaunchKernelGGL(kern, dim3(blocks_for((n - j - 1) * G, ROW_THREADS)), dim3(ROW_THREADS), lds,
                                   st, (uint32_t *)A, n, lda, lower, j, status);
        });
    });
    HIP_CHECK(hipGetLastError());
}

void mp_trsm_lower(int limbs, const void *A, int64_t n, void *B, int64_t nrhs, int transpose, int *fail,
                   hipStream_t st) {
    if (n == 0 || nrhs == 0) return;
    SDPX_DISPATCH_W(limbs, {
        SDPX_DISPATCH_G({
            auto kern = trsm_lower_kernel<W, G>;
            const size_t lds = panel_lds<W>((const void *)kern, ROW_THREADS);
            hipLaunchKernelGGL(kern, dim3(blocks_for(nrhs * G, ROW_THREADS)), dim3(ROW_THREADS), lds, st,
                               (const uint32_t *)A, n, (uint32_t *)B, nrhs, transpose, fail);
        });
    });
    HIP_CHECK(hipGetLastError());
}

void mp_potrs(int limbs, const void *A, int64_t n, int64_t lda, int lower, void *B, int64_t ldb,
              int64_t nrhs, hipStream_t st) {
    if (n == 0 || nrhs == 0) return;
    SDPX_DISPATCH_W(limbs, {
        hipLaunchKernelGGL(potrs_kernel<W>, dim3(blocks_for(nrhs, ROW_THREADS)), dim3(ROW_THREADS), 0, st,
                           (const uint32_t *)A, n, lda, lower, (uint32_t *)B, ldb, nrhs);
    });
    HIP_CHECK(hipGetLastError());
}

void mp_mirror_upper(int limbs, void *S, int64_t n, int64_t ld, hipStream_t st) {
    if (n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(mirror_upper_kernel<W>, dim3(blocks_for(n * n, 256)),
                                              dim3(256), 0, st, (uint32_t *)S, n, ld));
    HIP_CHECK(hipGetLastError());
}

void mp_mirror_lower(int limbs, void *S, int64_t n, int64_t ld, hipStream_t st) {
    if (n == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(mirror_lower_kernel<W>, dim3(blocks_for(n * n, 256)),
                                              dim3(256), 0, st, (uint32_t *)S, n, ld));
    HIP_CHECK(hipGetLastError());
}

void mp_col_absmax(int limbs, const void *X, int64_t ldx, int64_t m, int64_t ncols, void *out, hipStream_t st) {
    if (ncols == 0) return;
    SDPX_DISPATCH_W(limbs, hipLaunchKernelGGL(col_absmax_kernel<W>, dim3((unsigned)ncols), dim3(256), 0, st,
                                              (const uint32_t *)X, ldx, m, (uint32_t *)out));
    HIP_CHECK(hipGetLastError());
}

}  // namespace sdpx
