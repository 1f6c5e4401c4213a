/* CPU restatement of the SDPX exact MPFR product and LDL^T paths.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, by bench.py as the
 * timed CPU baseline, and by __graft_entry__.smoke(); never by the product.
 *
 * Number format: crates/arithmetic/src/lib.rs:333-340 (limbs LSB first,
 * MPFR kind code, i32 exponent; value = M * 2^(e - 64N)), zero as in
 * lib.rs:366-373.  Rounding: exact accumulation, one nearest-even rounding
 * (crates/arithmetic/src/exactdot.rs:1-13, from_scaled_integer 1316-1391).
 * Products: the contract of crates/solver/src/algebra/dense/blas/
 * rns_blas.rs:2134-2417 (decline on non-finite entries or an operand
 * exponent spread above 4096, zeros for a zero operand, packed upper
 * columns when square and upper-only).  LDL^T: arrow.rs:156-490
 * (ldl_dots, ldl_rec, solve_unit_lower_t) for dense envelopes.
 *
 * Independent of the GPU library: its own big-integer accumulator and its
 * own rounding code.  Checked against oracle/mpexact.py (plain Python
 * integers) and the reference's known answers in tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define KIND_NAN 0
#define KIND_INF 1
#define KIND_ZERO 2
#define KIND_REG 3
#define EMAX ((1 << 30) - 1)
#define EMIN (1 - (1 << 30))
#define MAX_SPREAD 4096

/* One value: limbs[0..n), then kind (i32), exponent (i32). */
typedef struct {
    int n;
    size_t stride; /* bytes per value: 8n + 8 */
} fmt_t;

static inline const uint64_t *limbs_of(const void *base, size_t stride, size_t i) {
    return (const uint64_t *)((const char *)base + i * stride);
}
static inline int32_t kind_of(const void *base, size_t stride, size_t i, int n) {
    return *(const int32_t *)((const char *)base + i * stride + 8 * (size_t)n);
}
static inline int32_t exp_of(const void *base, size_t stride, size_t i, int n) {
    return *(const int32_t *)((const char *)base + i * stride + 8 * (size_t)n + 4);
}

static void write_zero(void *base, size_t stride, size_t i, int n) {
    char *p = (char *)base + i * stride;
    memset(p, 0, 8 * (size_t)n);
    *(int32_t *)(p + 8 * n) = KIND_ZERO;
    *(int32_t *)(p + 8 * n + 4) = 0;
}

static int bitlen64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

/* Round (-1)^neg * mag * 2^scale once, nearest-even, to 64n bits and write
 * it at out[i] (exactdot.rs:1316-1391).  mag has len limbs, LSB first. */
static void from_scaled_integer(int neg, const uint64_t *mag, int len, int64_t scale, int n,
                                void *out, size_t stride, size_t i) {
    while (len > 0 && mag[len - 1] == 0) len--;
    if (len == 0) {
        write_zero(out, stride, i, n);
        return;
    }
    int64_t bits = (int64_t)(len - 1) * 64 + bitlen64(mag[len - 1]);
    int64_t p = 64 * (int64_t)n;
    int64_t shift = bits - p;
    uint64_t m[64];
    for (int l = 0; l < n; l++) {
        int64_t start = shift + 64 * (int64_t)l;
        uint64_t v = 0;
        /* floor division for negative starts */
        int64_t src = start >= 0 ? start / 64 : -((-start + 63) / 64);
        int64_t off = start - src * 64;
        if (src >= 0 && src < len) v |= mag[src] >> off;
        if (off != 0 && src + 1 >= 0 && src + 1 < len) v |= mag[src + 1] << (64 - off);
        m[l] = v;
    }
    if (shift > 0) {
        int64_t rb = shift - 1;
        int round = (mag[rb / 64] >> (rb % 64)) & 1;
        int sticky = 0;
        for (int64_t w = 0; w < rb / 64 && !sticky; w++) sticky = mag[w] != 0;
        if (!sticky && rb % 64) sticky = (mag[rb / 64] & ((1ull << (rb % 64)) - 1)) != 0;
        if (round && (sticky || (m[0] & 1))) {
            int carry = 1;
            for (int l = 0; l < n && carry; l++) {
                m[l] += 1;
                carry = m[l] == 0;
            }
            if (carry) {
                m[n - 1] = 1ull << 63;
                shift += 1;
            }
        }
    }
    int64_t e = scale + shift + p;
    char *dst = (char *)out + i * stride;
    if (e > EMAX) {
        memset(dst, 0, 8 * (size_t)n);
        *(int32_t *)(dst + 8 * n) = neg ? -KIND_INF : KIND_INF;
        *(int32_t *)(dst + 8 * n + 4) = 0;
        return;
    }
    if (e < EMIN) {
        write_zero(out, stride, i, n);
        return;
    }
    memcpy(dst, m, 8 * (size_t)n);
    *(int32_t *)(dst + 8 * n) = neg ? -KIND_REG : KIND_REG;
    *(int32_t *)(dst + 8 * n + 4) = (int32_t)e;
}

/* Add x (len limbs) << shift into acc (acc has room). */
static void add_shifted(uint64_t *acc, int acc_len, const uint64_t *x, int len, int64_t shift) {
    int64_t w = shift / 64;
    int off = (int)(shift % 64);
    u128 carry = 0;
    int64_t j = 0;
    for (; j <= len; j++) {
        uint64_t lo = j < len ? x[j] : 0;
        uint64_t prev = j > 0 ? x[j - 1] : 0;
        uint64_t v = off ? ((lo << off) | (prev >> (64 - off))) : lo;
        if (!off && j == len) v = 0;
        u128 s = (u128)acc[w + j] + v + carry;
        acc[w + j] = (uint64_t)s;
        carry = s >> 64;
    }
    for (int64_t t = w + j; carry && t < acc_len; t++) {
        u128 s = (u128)acc[t] + carry;
        acc[t] = (uint64_t)s;
        carry = s >> 64;
    }
}

/* Compare magnitudes, return sign of (a - b). */
static int cmp_mag(const uint64_t *a, const uint64_t *b, int len) {
    for (int i = len - 1; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
    }
    return 0;
}
static void sub_mag(uint64_t *a, const uint64_t *b, int len) { /* a -= b, a >= b */
    uint64_t borrow = 0;
    for (int i = 0; i < len; i++) {
        u128 d = (u128)a[i] - b[i] - borrow;
        a[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) ? 1 : 0;
    }
}

/* Exact dot of k pairs x[p], y[p] (element pointers given via index arrays),
 * rounded once into out[oi].  Non-finite inputs are not expected here. */
static void exact_dot(int n, int k, const void *xa, size_t xs, const size_t *xi, const void *ya,
                      size_t ys, const size_t *yi, void *out, size_t os, size_t oi) {
    size_t stride = 8 * (size_t)n + 8;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int p = 0; p < k; p++) {
        int32_t ka = kind_of(xa, stride, xi[p], n), kb = kind_of(ya, stride, yi[p], n);
        if (abs(ka) != KIND_REG || abs(kb) != KIND_REG) continue;
        int64_t s = (int64_t)exp_of(xa, stride, xi[p], n) + exp_of(ya, stride, yi[p], n) - 128 * (int64_t)n;
        if (s < lo) lo = s;
        if (s > hi) hi = s;
    }
    if (lo == INT64_MAX) {
        write_zero(out, os, oi, n);
        return;
    }
    int64_t width = hi - lo + 128 * (int64_t)n + 64 + 64;
    int acc_len = (int)(width / 64 + 2);
    uint64_t *pos = calloc((size_t)acc_len, 8), *neg = calloc((size_t)acc_len, 8);
    uint64_t prod[128];
    for (int p = 0; p < k; p++) {
        int32_t ka = kind_of(xa, stride, xi[p], n), kb = kind_of(ya, stride, yi[p], n);
        if (abs(ka) != KIND_REG || abs(kb) != KIND_REG) continue;
        const uint64_t *ma = limbs_of(xa, stride, xi[p]), *mb = limbs_of(ya, stride, yi[p]);
        memset(prod, 0, 16 * (size_t)n);
        for (int i = 0; i < n; i++) {
            u128 carry = 0;
            for (int j = 0; j < n; j++) {
                u128 t = (u128)ma[i] * mb[j] + prod[i + j] + carry;
                prod[i + j] = (uint64_t)t;
                carry = t >> 64;
            }
            prod[i + n] = (uint64_t)carry;
        }
        int64_t s = (int64_t)exp_of(xa, stride, xi[p], n) + exp_of(ya, stride, yi[p], n) - 128 * (int64_t)n;
        int negative = (ka < 0) != (kb < 0);
        add_shifted(negative ? neg : pos, acc_len, prod, 2 * n, s - lo);
    }
    int c = cmp_mag(pos, neg, acc_len);
    if (c >= 0) {
        sub_mag(pos, neg, acc_len);
        from_scaled_integer(0, pos, acc_len, lo, n, out, os, oi);
    } else {
        sub_mag(neg, pos, acc_len);
        from_scaled_integer(1, neg, acc_len, lo, n, out, os, oi);
    }
    free(pos);
    free(neg);
}

/* Exponent range of the stored entries of a (rows x cols, ld) operand.
 * Returns 0 with *lo > *hi for an all-zero operand, -1 for non-finite. */
static int exponent_range(int n, const void *a, size_t rows, size_t cols, size_t ld, int64_t *lo,
                          int64_t *hi) {
    size_t stride = 8 * (size_t)n + 8;
    *lo = INT64_MAX;
    *hi = INT64_MIN;
    for (size_t j = 0; j < cols; j++)
        for (size_t i = 0; i < rows; i++) {
            int32_t k = kind_of(a, stride, i + j * ld, n);
            if (abs(k) == KIND_ZERO) continue;
            if (abs(k) != KIND_REG) return -1;
            int64_t e = exp_of(a, stride, i + j * ld, n) - 64 * (int64_t)n;
            if (e < *lo) *lo = e;
            if (e > *hi) *hi = e;
        }
    return 0;
}

/* Exact product op(a) op(b) (rns_blas.rs:2134-2417 contract).
 * Returns 1 on success, 0 on decline (out unwritten).  Packed output when
 * upper_only, m == n and packed != 0; otherwise ldc = m and, when
 * upper_only, only i <= j is written.  threads > 1 uses OpenMP. */
int oracle_gemm_exact(int n, char ta, char tb, size_t m, size_t nn, size_t k, const void *a,
                      size_t lda, const void *b, size_t ldb, int upper_only, int packed, void *out,
                      int threads) {
    size_t stride = 8 * (size_t)n + 8;
    int tan = (ta == 'N' || ta == 'n'), tbn = (tb == 'N' || tb == 'n');
    int64_t la, ha, lb, hb;
    if (exponent_range(n, a, tan ? m : k, tan ? k : m, lda, &la, &ha) < 0) return 0;
    if (exponent_range(n, b, tbn ? k : nn, tbn ? nn : k, ldb, &lb, &hb) < 0) return 0;
    if ((la <= ha && ha - la > MAX_SPREAD) || (lb <= hb && hb - lb > MAX_SPREAD)) return 0;
    packed = packed && upper_only && m == nn;
    long long total = (long long)nn;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (long long jj = 0; jj < total; jj++) {
        size_t j = (size_t)jj;
        size_t *xi = malloc(sizeof(size_t) * (k ? k : 1));
        size_t *yi = malloc(sizeof(size_t) * (k ? k : 1));
        for (size_t p = 0; p < k; p++) yi[p] = tbn ? p + j * ldb : j + p * ldb;
        size_t iend = upper_only ? (j + 1 < m ? j + 1 : m) : m;
        for (size_t i = 0; i < iend; i++) {
            for (size_t p = 0; p < k; p++) xi[p] = tan ? i + p * lda : p + i * lda;
            size_t o = packed ? j * (j + 1) / 2 + i : i + j * m;
            exact_dot(n, (int)k, a, stride, xi, b, stride, yi, out, stride, o);
        }
        free(xi);
        free(yi);
    }
    return 1;
}

/* ------------------------------------------------------------------------
 * The reference's CPU exact product restated: residues over primes of 18-26
 * bits computed in binary64 (rns_blas.rs: primes 95-112, prime_bits
 * 184-188, Plan 705-737, the gemm contract 2134-2417), for the CPU
 * baseline, taken where the reference takes it
 * (residue_product_profitable, blas/mpfr.rs:349-361).  Per operand one
 * window (the lowest set weight), each value an integer X = M 2^s there;
 * the primes (descending below 2^b, b = floor((55 - ceil log2 k) / 2) in
 * [18, 26]) multiply to at least 2^(wa + wb + log2 k + 3) for operand
 * widths wa, wb; residues |r| < p/2; per prime a binary64 product exact
 * over inner chunks of 2^(55 - 2b) terms, reduced mod p; then the CRT
 * S = sum x_q (M/q) - round(sum x_q / q) M with x_q = c_q (M/q)^-1 mod q
 * (the reference's crt: signed lift with an estimate kept inside
 * (-1/4, 1/4) by M > 4|S|), and one rounding.  Same bits as the exact dots
 * by construction; tests check it.
 * ---------------------------------------------------------------------- */
static int rns_products = 0;
void oracle_set_rns_products(int on) { rns_products = on; }

static int is_prime_small(uint64_t v) {
    if (v < 2) return 0;
    for (uint64_t d = 2; d * d <= v; d++)
        if (v % d == 0) return 0;
    return 1;
}

static uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = (uint64_t)((u128)r * b % q);
        b = (uint64_t)((u128)b * b % q);
        e >>= 1;
    }
    return r;
}

/* residue of the value at index idx, as an integer at window lo, mod q, in
 * (-q/2, q/2); pw[s] = 2^s mod q for s up to the operand's width */
static double rns_residue(int n, const void *base, size_t stride, size_t idx, int64_t lo, uint64_t q,
                          const uint64_t *pw) {
    int32_t k = kind_of(base, stride, idx, n);
    if (abs(k) != KIND_REG) return 0.0;
    const uint64_t *l = limbs_of(base, stride, idx);
    int64_t sh = (int64_t)exp_of(base, stride, idx, n) - 64 * (int64_t)n - lo;
    uint64_t r = 0;
    for (int i = 0; i < n; i++) r += (uint64_t)((u128)(l[i] % q) * pw[sh + 64 * i] % q);
    r %= q;
    if (k < 0 && r) r = q - r;
    return r > (q - 1) / 2 ? (double)r - (double)q : (double)r;
}

/* x mod q in [0, q) for an integer-valued x with |x| < 2^53 */
static inline double mod_exact(double x, double q, double qinv) {
    double r = x - q * nearbyint(x * qinv);
    if (r < 0) r += q;
    if (r >= q) r -= q;
    return r;
}

static uint64_t inv_mod(uint64_t a, uint64_t q) { return powmod_u64(a, q - 2, q); }

int oracle_gemm_rns(int n, char ta, char tb, size_t m, size_t nn, size_t k, const void *a, size_t lda,
                    const void *b, size_t ldb, int upper_only, int packed, void *out, int threads) {
    size_t stride = 8 * (size_t)n + 8;
    int tan = (ta == 'N' || ta == 'n'), tbn = (tb == 'N' || tb == 'n');
    int64_t la, ha, lb, hb;
    if (exponent_range(n, a, tan ? m : k, tan ? k : m, lda, &la, &ha) < 0) return 0;
    if (exponent_range(n, b, tbn ? k : nn, tbn ? nn : k, ldb, &lb, &hb) < 0) return 0;
    if ((la <= ha && ha - la > MAX_SPREAD) || (lb <= hb && hb - lb > MAX_SPREAD)) return 0;
    packed = packed && upper_only && m == nn;
    const int64_t P = 64 * (int64_t)n;
    if (la > ha || lb > hb || k == 0) { /* a zero operand: zeros */
        for (size_t j = 0; j < nn; j++)
            for (size_t i = 0; i < (upper_only ? (j + 1 < m ? j + 1 : m) : m); i++)
                write_zero(out, stride, packed ? j * (j + 1) / 2 + i : i + j * m, n);
        return 1;
    }
    int lk = 0;
    while (((size_t)1 << lk) < k) lk++;
    int bits = (int)((53.0 + 2.0 - lk) / 2.0);
    if (bits < 18) bits = 18;
    if (bits > 26) bits = 26;
    const double needed = (double)(P + ha - la) + (double)(P + hb - lb) + log2((double)k) + 3.0;
    uint64_t qs[1024];
    int L = 0;
    double have = 0.0;
    for (uint64_t v = (1ull << bits) - 1; have < needed; v -= 2) {
        if (!is_prime_small(v)) continue;
        if (L == 1024) return 0;
        qs[L++] = v;
        have += log2((double)v);
    }
    const size_t kc = (size_t)1 << (53 + 2 - 2 * bits < 40 ? 53 + 2 - 2 * bits : 40);
    /* residues: A rows (i, p) at i k + p, B columns (p, j) at j k + p */
    double *ra = malloc(sizeof(double) * (size_t)L * m * k), *rb = malloc(sizeof(double) * (size_t)L * nn * k);
    const size_t npw = (size_t)(P + (ha - la > hb - lb ? ha - la : hb - lb)) + 64;
    uint64_t *pw = malloc(8 * (size_t)L * npw);
    for (int q = 0; q < L; q++) {
        uint64_t *t = pw + (size_t)q * npw;
        t[0] = 1 % qs[q];
        for (size_t e = 1; e < npw; e++) t[e] = (t[e - 1] * 2) % qs[q];
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (long long row = 0; row < (long long)(m + nn); row++) {
        for (int q = 0; q < L; q++) {
            const uint64_t *t = pw + (size_t)q * npw;
            if ((size_t)row < m) {
                size_t i = (size_t)row;
                for (size_t p = 0; p < k; p++)
                    ra[((size_t)q * m + i) * k + p] =
                        rns_residue(n, a, stride, tan ? i + p * lda : p + i * lda, la, qs[q], t);
            } else {
                size_t j = (size_t)row - m;
                for (size_t p = 0; p < k; p++)
                    rb[((size_t)q * nn + j) * k + p] =
                        rns_residue(n, b, stride, tbn ? p + j * ldb : j + p * ldb, lb, qs[q], t);
            }
        }
    }
    free(pw);
    /* M, M/q (limbs) and the CRT weights (M/q)^-1 mod q */
    int lm = (int)(have / 64.0) + 2;
    uint64_t *M = calloc((size_t)lm, 8), *Mq = calloc((size_t)L * lm, 8), *w = malloc(8 * (size_t)L);
    M[0] = 1;
    for (int q = 0; q < L; q++) {
        u128 c = 0;
        for (int i = 0; i < lm; i++) {
            c += (u128)M[i] * qs[q];
            M[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    for (int q = 0; q < L; q++) {
        uint64_t *d = Mq + (size_t)q * lm;
        d[0] = 1;
        uint64_t r = 1; /* (M/q) mod q */
        for (int o = 0; o < L; o++) {
            if (o == q) continue;
            u128 c = 0;
            for (int i = 0; i < lm; i++) {
                c += (u128)d[i] * qs[o];
                d[i] = (uint64_t)c;
                c >>= 64;
            }
            r = (uint64_t)((u128)r * (qs[o] % qs[q]) % qs[q]);
        }
        w[q] = inv_mod(r, qs[q]);
    }
    long long total = (long long)nn;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (long long jj = 0; jj < total; jj++) {
        size_t j = (size_t)jj;
        uint64_t *acc = malloc(8 * (size_t)(lm + 1));
        uint32_t *cres = malloc(4 * (size_t)L);
        size_t iend = upper_only ? (j + 1 < m ? j + 1 : m) : m;
        for (size_t i = 0; i < iend; i++) {
            for (int q = 0; q < L; q++) {
                const double *x = ra + ((size_t)q * m + i) * k, *y = rb + ((size_t)q * nn + j) * k;
                const double qd = (double)qs[q], qinv = 1.0 / qd;
                double r = 0.0;
                for (size_t p0 = 0; p0 < k; p0 += kc) {
                    size_t p1 = p0 + kc < k ? p0 + kc : k;
                    double s = 0.0;
                    /* every partial sum is an integer below 2^53: any order gives the same value */
#pragma omp simd reduction(+ : s)
                    for (size_t p = p0; p < p1; p++) s += x[p] * y[p];
                    r = mod_exact(r + mod_exact(s, qd, qinv), qd, qinv);
                }
                cres[q] = (uint32_t)r;
            }
            /* S = sum x_q (M/q) - K M, K = round(sum x_q / q) */
            memset(acc, 0, 8 * (size_t)(lm + 1));
            double frac = 0.0;
            for (int q = 0; q < L; q++) {
                uint64_t xq = (uint64_t)((u128)cres[q] * w[q] % qs[q]);
                frac += (double)xq / (double)qs[q];
                const uint64_t *d = Mq + (size_t)q * lm;
                u128 c = 0;
                for (int t = 0; t < lm; t++) {
                    c += (u128)d[t] * xq + acc[t];
                    acc[t] = (uint64_t)c;
                    c >>= 64;
                }
                acc[lm] += (uint64_t)c;
            }
            uint64_t K = (uint64_t)llround(frac);
            u128 bor = 0;
            {
                u128 c = 0;
                for (int t = 0; t <= lm; t++) {
                    c += (u128)(t < lm ? M[t] : 0) * K;
                    uint64_t sub = (uint64_t)c;
                    c >>= 64;
                    u128 dd = (u128)acc[t] - sub - bor;
                    acc[t] = (uint64_t)dd;
                    bor = (dd >> 64) ? 1 : 0;
                }
            }
            int neg = (acc[lm] >> 63) != 0;
            if (neg) {
                int carry = 1;
                for (int t = 0; t <= lm; t++) {
                    acc[t] = ~acc[t];
                    if (carry) {
                        acc[t] += 1;
                        carry = acc[t] == 0;
                    }
                }
            }
            from_scaled_integer(neg, acc, lm + 1, la + lb, n, out, stride,
                                packed ? j * (j + 1) / 2 + i : i + j * m);
        }
        free(acc);
        free(cres);
    }
    free(ra);
    free(rb);
    free(M);
    free(Mq);
    free(w);
    return 1;
}

/* The product the restatements use: the CPU residue product where the
 * reference takes it (when enabled for the CPU baseline), else exact dots. */
static int ora_gemm(int n, char ta, char tb, size_t m, size_t nn, size_t k, const void *a, size_t lda,
                    const void *b, size_t ldb, int upper_only, int packed, void *out, int threads) {
    int profitable = (n >= 4 && k >= 24 && m * nn >= 576) || (n >= 2 && k >= 64 && m * nn >= 4096);
    if (rns_products && profitable &&
        oracle_gemm_rns(n, ta, tb, m, nn, k, a, lda, b, ldb, upper_only, packed, out, threads))
        return 1;
    return oracle_gemm_exact(n, ta, tb, m, nn, k, a, lda, b, ldb, upper_only, packed, out, threads);
}

/* Exact dots of an explicit list of outputs (spot checks of large products):
 * out[t] = exact op(a)(rows[t], :) . op(b)(:, cols[t]). */
void oracle_gemm_entries(int n, char ta, char tb, size_t k, const void *a, size_t lda,
                         const void *b, size_t ldb, size_t count, const int64_t *rows,
                         const int64_t *cols, void *out, int threads) {
    size_t stride = 8 * (size_t)n + 8;
    int tan = (ta == 'N' || ta == 'n'), tbn = (tb == 'N' || tb == 'n');
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
    for (long long t = 0; t < (long long)count; t++) {
        size_t i = (size_t)rows[t], j = (size_t)cols[t];
        size_t *xi = malloc(sizeof(size_t) * (k ? k : 1));
        size_t *yi = malloc(sizeof(size_t) * (k ? k : 1));
        for (size_t p = 0; p < k; p++) {
            xi[p] = tan ? i + p * lda : p + i * lda;
            yi[p] = tbn ? p + j * ldb : j + p * ldb;
        }
        exact_dot(n, (int)k, a, stride, xi, b, stride, yi, out, stride, (size_t)t);
        free(xi);
        free(yi);
    }
}

/* ------------------------------------------------------------------------
 * Scalar operations, each one exact value rounded once.
 * ---------------------------------------------------------------------- */
typedef struct {
    int n;
    size_t stride;
    uint8_t one[8 * 32 + 8], minus_one[8 * 32 + 8];
} ora_ctx;

static void ora_init(ora_ctx *c, int n) {
    c->n = n;
    c->stride = 8 * (size_t)n + 8;
    memset(c->one, 0, sizeof c->one);
    ((uint64_t *)c->one)[n - 1] = 1ull << 63;
    *(int32_t *)(c->one + 8 * n) = KIND_REG;
    *(int32_t *)(c->one + 8 * n + 4) = 1;
    memcpy(c->minus_one, c->one, c->stride);
    *(int32_t *)(c->minus_one + 8 * n) = -KIND_REG;
}

/* out = RN(sum_i x[i] * y[i]) over pointer lists */
static void ora_dot_ptrs(const ora_ctx *c, int k, const uint8_t **x, const uint8_t **y, uint8_t *out) {
    /* gather into a contiguous pair of temporaries for exact_dot */
    size_t st = c->stride;
    uint8_t *xa = malloc(st * (k ? k : 1)), *ya = malloc(st * (k ? k : 1));
    size_t *idx = malloc(sizeof(size_t) * (k ? k : 1));
    for (int p = 0; p < k; p++) {
        memcpy(xa + p * st, x[p], st);
        memcpy(ya + p * st, y[p], st);
        idx[p] = (size_t)p;
    }
    exact_dot(c->n, k, xa, st, idx, ya, st, idx, out, st, 0);
    free(xa);
    free(ya);
    free(idx);
}

static void ora_sub(const ora_ctx *c, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    const uint8_t *x[2] = {a, b}, *y[2] = {c->one, c->minus_one};
    ora_dot_ptrs(c, 2, x, y, out);
}
static void ora_mul(const ora_ctx *c, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    const uint8_t *x[1] = {a}, *y[1] = {b};
    ora_dot_ptrs(c, 1, x, y, out);
}
static int ora_is_zero(const ora_ctx *c, const uint8_t *v) { return abs(*(const int32_t *)(v + 8 * c->n)) == KIND_ZERO; }
static int ora_is_finite(const ora_ctx *c, const uint8_t *v) {
    int k = abs(*(const int32_t *)(v + 8 * c->n));
    return k == KIND_ZERO || k == KIND_REG;
}
static int ora_is_neg(const ora_ctx *c, const uint8_t *v) { return *(const int32_t *)(v + 8 * c->n) < 0; }
static void ora_neg(const ora_ctx *c, uint8_t *v) {
    int32_t *k = (int32_t *)(v + 8 * c->n);
    if (abs(*k) == KIND_REG || abs(*k) == KIND_INF) *k = -*k;
}

/* RN(1/d): quotient digits of 2^(2P+3) / M by shift-and-subtract over
 * 32-bit words, remainder as the sticky bit. */
static void ora_recip(const ora_ctx *c, const uint8_t *d, uint8_t *out) {
    int n = c->n, P = 64 * n;
    const uint64_t *m = (const uint64_t *)d;
    int32_t e = *(const int32_t *)(d + 8 * n + 4);
    int words = 2 * n + 2; /* remainder words (32-bit) */
    uint32_t rem[70] = {0}, mw[70] = {0};
    for (int i = 0; i < n; i++) {
        mw[2 * i] = (uint32_t)m[i];
        mw[2 * i + 1] = (uint32_t)(m[i] >> 32);
    }
    int qbits = P + 4;
    uint64_t q[40] = {0};
    /* numerator 2^(2P+3): bits 2P+3 .. 0; quotient bits below qbits only */
    for (int bit = 2 * P + 3; bit >= 0; bit--) {
        uint32_t carry = (bit == 2 * P + 3);
        for (int i = 0; i < words; i++) {
            uint32_t v = rem[i];
            rem[i] = (v << 1) | carry;
            carry = v >> 31;
        }
        int ge = 1;
        for (int i = words - 1; i >= 0; i--) {
            if (rem[i] != mw[i]) {
                ge = rem[i] > mw[i];
                break;
            }
        }
        if (ge) {
            int64_t br = 0;
            for (int i = 0; i < words; i++) {
                int64_t t = (int64_t)rem[i] - mw[i] - br;
                br = t < 0;
                rem[i] = (uint32_t)t;
            }
            if (bit < qbits + 2) q[bit / 64] |= 1ull << (bit % 64);
        }
    }
    int sticky = 0;
    for (int i = 0; i < words; i++) sticky |= rem[i] != 0;
    /* 1/d = 2^(P-e)/M = (2^(2P+3)/M) 2^(-P-3-e); append sticky below */
    uint64_t mag[41] = {0};
    uint64_t carry = (uint64_t)sticky;
    for (int i = 0; i < 40; i++) {
        mag[i] = (q[i] << 1) | carry;
        carry = q[i] >> 63;
    }
    mag[40] = carry;
    from_scaled_integer(ora_is_neg(c, d), mag, 41, -(int64_t)P - 3 - e - 1, n, out, c->stride, 0);
}

/* Little-endian 32-bit word helpers for the division and square root. */
#define BW 200 /* words: radicands and numerators up to 2P + 8 bits, P <= 2048 */
static void bw_from_mant(uint32_t *w, const uint8_t *v, int n) {
    memset(w, 0, sizeof(uint32_t) * BW);
    const uint64_t *m = (const uint64_t *)v;
    for (int i = 0; i < n; i++) {
        w[2 * i] = (uint32_t)m[i];
        w[2 * i + 1] = (uint32_t)(m[i] >> 32);
    }
}
static int bw_bit(const uint32_t *w, int b) { return (w[b >> 5] >> (b & 31)) & 1; }
static void bw_shl1(uint32_t *w, int words, uint32_t in) {
    for (int i = 0; i < words; i++) {
        uint32_t v = w[i];
        w[i] = (v << 1) | in;
        in = v >> 31;
    }
}
static int bw_ge(const uint32_t *a, const uint32_t *b, int words) {
    for (int i = words - 1; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i];
    return 1;
}
static void bw_sub(uint32_t *a, const uint32_t *b, int words) {
    int64_t br = 0;
    for (int i = 0; i < words; i++) {
        int64_t t = (int64_t)a[i] - b[i] - br;
        br = t < 0;
        a[i] = (uint32_t)t;
    }
}
static int bw_nonzero(const uint32_t *a, int words) {
    for (int i = 0; i < words; i++)
        if (a[i]) return 1;
    return 0;
}
static void bw_round(const ora_ctx *c, int neg, const uint32_t *q, int qwords, int sticky,
                     int64_t scale, uint8_t *out) {
    /* (q << 1 | sticky) * 2^(scale - 1), one rounding */
    uint64_t mag[BW / 2 + 2] = {0};
    uint32_t w[BW + 2] = {0};
    uint32_t carry = (uint32_t)sticky;
    for (int i = 0; i < qwords; i++) {
        w[i] = (q[i] << 1) | carry;
        carry = q[i] >> 31;
    }
    w[qwords] = carry;
    int words = qwords + 1;
    for (int i = 0; i < words; i += 2) mag[i / 2] = w[i] | ((uint64_t)w[i + 1] << 32);
    from_scaled_integer(neg, mag, (words + 1) / 2, scale - 1, c->n, out, c->stride, 0);
}

/* RN(x / d): restoring division of Mx 2^(P+2) by Md, remainder as sticky. */
static void ora_div(const ora_ctx *c, const uint8_t *x, const uint8_t *d, uint8_t *out) {
    int n = c->n, P = 64 * n;
    if (ora_is_zero(c, x)) {
        memset(out, 0, c->stride);
        *(int32_t *)(out + 8 * n) = KIND_ZERO;
        return;
    }
    uint32_t num[BW], den[BW], rem[BW] = {0}, q[BW] = {0};
    bw_from_mant(num, x, n);
    bw_from_mant(den, d, n);
    int words = 2 * n + 2;
    int nbits = P + (P + 2); /* numerator Mx << (P + 2) */
    for (int b = nbits - 1; b >= 0; b--) {
        int src = b - (P + 2);
        bw_shl1(rem, words, src >= 0 ? (uint32_t)bw_bit(num, src) : 0u);
        bw_shl1(q, words, 0);
        if (bw_ge(rem, den, words)) {
            bw_sub(rem, den, words);
            q[0] |= 1;
        }
    }
    int32_t ex = *(const int32_t *)(x + 8 * n + 4), ed = *(const int32_t *)(d + 8 * n + 4);
    bw_round(c, ora_is_neg(c, x) != ora_is_neg(c, d), q, words, bw_nonzero(rem, words),
             (int64_t)ex - ed - P - 2, out);
}

/* RN(sqrt(a)), a > 0: restoring integer square root (two bits a step) of
 * the mantissa scaled to an even exponent and 2P + 4 or more bits. */
static void ora_sqrt(const ora_ctx *c, const uint8_t *a, uint8_t *out) {
    int n = c->n, P = 64 * n;
    if (ora_is_zero(c, a)) {
        memset(out, 0, c->stride);
        *(int32_t *)(out + 8 * n) = KIND_ZERO;
        return;
    }
    uint32_t m[BW];
    bw_from_mant(m, a, n);
    int64_t e = (int64_t)*(const int32_t *)(a + 8 * n + 4) - P; /* value = M 2^e */
    int extra = 0;                                               /* radicand = M << extra */
    if (e & 1) {
        extra = 1;
        e -= 1;
    }
    int k = (2 * P + 4 - (P + extra) + 1) / 2;
    if (k < 0) k = 0;
    extra += 2 * k;
    int rbits = P + extra; /* radicand bits (even count after padding) */
    if (rbits & 1) rbits++;
    int words = (rbits / 2 + 40) / 32 + 2;
    uint32_t rem[BW] = {0}, root[BW] = {0}, trial[BW];
    for (int b = rbits - 2; b >= 0; b -= 2) {
        int s1 = b + 1 - extra, s0 = b - extra;
        uint32_t hi = s1 >= 0 && s1 < P ? (uint32_t)bw_bit(m, s1) : 0u;
        uint32_t lo = s0 >= 0 && s0 < P ? (uint32_t)bw_bit(m, s0) : 0u;
        bw_shl1(rem, words, hi);
        bw_shl1(rem, words, lo);
        memcpy(trial, root, sizeof(uint32_t) * words);
        bw_shl1(trial, words, 0);
        bw_shl1(trial, words, 1); /* (root << 2) | 1 */
        bw_shl1(root, words, 0);
        if (bw_ge(rem, trial, words)) {
            bw_sub(rem, trial, words);
            root[0] |= 1;
        }
    }
    bw_round(c, 0, root, words, bw_nonzero(rem, words), e / 2 - k, out);
}

/* RN(v - a b), one rounding (mul_add of -a, b, v) */
static void ora_fms(const ora_ctx *c, const uint8_t *v, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    uint8_t na[8 * 32 + 8];
    memcpy(na, a, c->stride);
    ora_neg(c, na);
    const uint8_t *x[2] = {v, na}, *y[2] = {c->one, b};
    ora_dot_ptrs(c, 2, x, y, out);
}

/* ------------------------------------------------------------------------
 * LDL^T restatement (arrow.rs:156-490): ldl_dots, solve_unit_lower_t,
 * ldl_rec with envelope cuts.  Products are exact dots rounded once.
 * ---------------------------------------------------------------------- */
#define AT(base, idx) ((base) + (size_t)(idx) * c->stride)
#define TRI(i) ((size_t)(i) * ((size_t)(i) + 1) / 2)

typedef struct {
    int reg;
    const uint8_t *eps, *delta;
    const int8_t *signs;
    size_t nsigns, offset;
    long long regularized;
    int threads;
} ora_ldl_opts;

static int ora_sign(const ora_ldl_opts *o, size_t k) {
    if (o->nsigns == 0) return 1;
    return o->signs[o->nsigns == 1 ? 0 : o->offset + k];
}

static int ora_pivot(const ora_ctx *c, uint8_t *d, int sign, ora_ldl_opts *o) {
    if (!ora_is_finite(c, d)) return 0;
    if (o->reg) {
        uint8_t t[8 * 32 + 8], diff[8 * 32 + 8];
        memcpy(t, d, c->stride);
        if (sign < 0) ora_neg(c, t);
        ora_sub(c, t, o->eps, diff);
        if (!ora_is_zero(c, diff) && ora_is_neg(c, diff)) {
            memcpy(d, o->delta, c->stride);
            if (sign < 0) ora_neg(c, d);
            o->regularized++;
        }
    }
    return !ora_is_zero(c, d);
}

static int ora_ldl_dots(const ora_ctx *c, uint8_t *l, size_t ld, size_t n, const size_t *first,
                        ora_ldl_opts *o, uint8_t *dinv) {
    size_t st = c->stride;
    uint8_t *r = calloc(TRI(n) + 1, st), *w = calloc(TRI(n) + 1, st), *d = calloc(n + 1, st);
    int ok = 1;
    for (size_t j = 0; j < n && ok; j++) {
        size_t fj = first[j];
        size_t *idx = malloc(sizeof(size_t) * (j + 1));
        for (size_t q = 0; q < j; q++) idx[q] = TRI(j) + q;
        uint8_t t[8 * 32 + 8];
        exact_dot(c->n, (int)(j - fj), r, st, idx + fj, w, st, idx + fj, t, st, 0);
        uint8_t *dj = AT(d, j);
        ora_sub(c, AT(l, j + j * ld), t, dj);
        if (ora_is_zero(c, dj) && j > fj) {
            /* exact a_jj - sum, rounded once */
            size_t k = j - fj + 1;
            const uint8_t **xs = malloc(sizeof(uint8_t *) * k), **ys = malloc(sizeof(uint8_t *) * k);
            for (size_t q = 0; q < j - fj; q++) {
                xs[q] = AT(r, TRI(j) + fj + q);
                ys[q] = AT(w, TRI(j) + fj + q);
            }
            xs[k - 1] = AT(l, j + j * ld);
            ys[k - 1] = c->minus_one;
            ora_dot_ptrs(c, (int)k, xs, ys, dj);
            ora_neg(c, dj);
            free(xs);
            free(ys);
        }
        free(idx);
        if (!ora_pivot(c, dj, ora_sign(o, j), o)) {
            ok = 0;
            break;
        }
        ora_recip(c, dj, AT(dinv, j));
#pragma omp parallel for schedule(dynamic, 8) num_threads(o->threads > 0 ? o->threads : 1)
        for (long long ii = (long long)j + 1; ii < (long long)n; ii++) {
            size_t i = (size_t)ii;
            uint8_t *rij = AT(r, TRI(i) + j);
            if (first[i] > j) {
                write_zero(rij, st, 0, c->n);
            } else {
                size_t lo = first[i] > first[j] ? first[i] : first[j];
                size_t *xi = malloc(sizeof(size_t) * (j + 1)), *yi = malloc(sizeof(size_t) * (j + 1));
                for (size_t q = lo; q < j; q++) {
                    xi[q - lo] = TRI(i) + q;
                    yi[q - lo] = TRI(j) + q;
                }
                uint8_t t2[8 * 32 + 8], diff[8 * 32 + 8];
                exact_dot(c->n, (int)(j - lo), r, st, xi, w, st, yi, t2, st, 0);
                ora_sub(c, AT(l, i + j * ld), t2, diff);
                ora_mul(c, diff, AT(dinv, j), rij);
                free(xi);
                free(yi);
            }
            ora_mul(c, rij, dj, AT(w, TRI(i) + j));
        }
    }
    if (ok) {
        for (size_t j = 0; j < n; j++) {
            memcpy(AT(l, j + j * ld), c->one, st);
            for (size_t i = j + 1; i < n; i++) memcpy(AT(l, i + j * ld), AT(r, TRI(i) + j), st);
        }
    }
    free(r);
    free(w);
    free(d);
    return ok;
}

/* c_out (m x n, ld m) = A (m x k, lda) * B^T where B is n x k (ldb). */
static void ora_product_nt(const ora_ctx *c, size_t m, size_t n, size_t k, const uint8_t *a,
                           size_t lda, const uint8_t *b, size_t ldb, uint8_t *out, int upper_packed,
                           int threads) {
    ora_gemm(c->n, 'N', 'T', m, n, k, a, lda, b, ldb, upper_packed, upper_packed, out, threads);
}

static void ora_solve_unit_lower_t(const ora_ctx *c, size_t m, size_t n, const uint8_t *L, size_t ldl,
                                   uint8_t *B, size_t ldb, const size_t *low, int threads) {
    size_t st = c->stride;
    if (m == 0 || n == 0) return;
    if (n <= 128) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
        for (long long ii = 0; ii < (long long)m; ii++) {
            size_t i = (size_t)ii;
            uint8_t *x = malloc(st * n);
            for (size_t j = 0; j < n; j++) memcpy(x + j * st, AT(B, i + j * ldb), st);
            size_t lo = low[i];
            size_t *xi = malloc(sizeof(size_t) * (n + 1)), *yi = malloc(sizeof(size_t) * (n + 1));
            for (size_t j = lo; j < n; j++) {
                for (size_t q = lo; q < j; q++) {
                    xi[q - lo] = q;
                    yi[q - lo] = j + q * ldl;
                }
                uint8_t t[8 * 32 + 8], u[8 * 32 + 8];
                exact_dot(c->n, (int)(j - lo), x, st, xi, L, st, yi, t, st, 0);
                ora_sub(c, x + j * st, t, u);
                memcpy(x + j * st, u, st);
            }
            for (size_t j = 0; j < n; j++) memcpy(AT(B, i + j * ldb), x + j * st, st);
            free(x);
            free(xi);
            free(yi);
        }
        return;
    }
    size_t n1 = n / 2, n2 = n - n / 2;
    size_t *low1 = malloc(sizeof(size_t) * m), *low2 = malloc(sizeof(size_t) * m);
    for (size_t i = 0; i < m; i++) {
        low1[i] = low[i] < n1 ? low[i] : n1;
        low2[i] = low[i] > n1 ? low[i] - n1 : 0;
    }
    ora_solve_unit_lower_t(c, m, n1, L, ldl, B, ldb, low1, threads);
    uint8_t *p = malloc(st * m * n2);
    ora_product_nt(c, m, n2, n1, B, ldb, AT(L, n1), ldl, p, 0, threads);
    uint8_t *B2 = AT(B, n1 * ldb);
    for (size_t j = 0; j < n2; j++)
        for (size_t i = 0; i < m; i++) {
            uint8_t t[8 * 32 + 8];
            ora_sub(c, AT(B2, i + j * ldb), AT(p, i + j * m), t);
            memcpy(AT(B2, i + j * ldb), t, st);
        }
    free(p);
    ora_solve_unit_lower_t(c, m, n2, AT(L, n1 + n1 * ldl), ldl, B2, ldb, low2, threads);
    free(low1);
    free(low2);
}

static int ora_ldl_rec(const ora_ctx *c, uint8_t *l, size_t ld, size_t n, const size_t *first,
                       ora_ldl_opts *o, uint8_t *dinv) {
    size_t st = c->stride;
    if (n <= 128) return ora_ldl_dots(c, l, ld, n, first, o, dinv);
    size_t low = n, cut = 0;
    int have = 0;
    for (size_t cc = n - 1; cc >= 1; cc--) {
        if (first[cc] < low) low = first[cc];
        size_t dc = cc > n / 2 ? cc - n / 2 : n / 2 - cc;
        size_t dk = cut > n / 2 ? cut - n / 2 : n / 2 - cut;
        if (low >= cc && (!have || dc < dk)) {
            cut = cc;
            have = 1;
        }
    }
    if (have) {
        size_t *tail = malloc(sizeof(size_t) * (n - cut));
        for (size_t i = cut; i < n; i++) tail[i - cut] = first[i] - cut;
        int ok1 = ora_ldl_rec(c, l, ld, cut, first, o, dinv);
        size_t save = o->offset;
        o->offset = save + cut;
        int ok2 = ora_ldl_rec(c, AT(l, cut + cut * ld), ld, n - cut, tail, o, AT(dinv, cut));
        o->offset = save;
        free(tail);
        return ok1 && ok2;
    }
    size_t n1 = n / 2, n2 = n - n / 2;
    if (!ora_ldl_rec(c, l, ld, n1, first, o, dinv)) return 0;
    uint8_t *l11 = malloc(st * n1 * n1);
    for (size_t j = 0; j < n1; j++) memcpy(l11 + j * n1 * st, AT(l, j * ld), n1 * st);
    size_t *lowv = malloc(sizeof(size_t) * n2);
    for (size_t i = 0; i < n2; i++) lowv[i] = first[n1 + i] < n1 ? first[n1 + i] : n1;
    uint8_t *a21 = AT(l, n1);
    ora_solve_unit_lower_t(c, n2, n1, l11, n1, a21, ld, lowv, o->threads);
    free(l11);
    free(lowv);
    uint8_t *w = malloc(st * n2 * n1);
    for (size_t j = 0; j < n1; j++) memcpy(w + j * n2 * st, AT(a21, j * ld), n2 * st);
    for (size_t j = 0; j < n1; j++)
        for (size_t i = 0; i < n2; i++) {
            uint8_t t[8 * 32 + 8];
            ora_mul(c, AT(a21, i + j * ld), AT(dinv, j), t);
            memcpy(AT(a21, i + j * ld), t, st);
        }
    uint8_t *packed = malloc(st * (n2 * (n2 + 1) / 2));
    ora_product_nt(c, n2, n2, n1, w, n2, a21, ld, packed, 1, o->threads);
    free(w);
    uint8_t *a22 = AT(l, n1 + n1 * ld);
    for (size_t a = 0; a < n2; a++)
        for (size_t b = a; b < n2; b++) {
            uint8_t t[8 * 32 + 8];
            ora_sub(c, AT(a22, b + a * ld), AT(packed, b * (b + 1) / 2 + a), t);
            memcpy(AT(a22, b + a * ld), t, st);
        }
    free(packed);
    size_t *tail = malloc(sizeof(size_t) * n2);
    for (size_t i = 0; i < n2; i++) tail[i] = first[n1 + i] > n1 ? first[n1 + i] - n1 : 0;
    size_t save = o->offset;
    o->offset = save + n1;
    int ok = ora_ldl_rec(c, a22, ld, n2, tail, o, AT(dinv, n1));
    o->offset = save;
    free(tail);
    return ok;
}

/* Entry point: factor the lower triangle of l (n x n, ld) in place.
 * mode 0: ldl_dots; 1: ldl_rec; 2: DenseLeaf::factor_loaded for wide types
 * (arrow.rs:808-850): from order REC_FACTOR_MIN = 256 (arrow.rs:62) ldl_rec,
 * and when it fails (one rounding per level can cancel a pivot to an exact
 * zero that exact dots keep nonzero, arrow.rs:820-846) the input is restored
 * and factored by ldl_dots; the failed attempt's regularization count is
 * dropped, as the reference's local count is; below order 256 ldl_dots.
 * The envelope is taken once, from the input.  reg_eps/reg_delta NULL: no
 * regularization.  Returns 1 on success; *regularized counts replaced
 * pivots; *retried (optional) is set when the exact-dots retry ran. */
int oracle_ldl_mode(int nl, uint8_t *l, size_t ld, size_t n, const int8_t *signs, size_t nsigns,
                    const uint8_t *reg_eps, const uint8_t *reg_delta, uint8_t *dinv, int mode,
                    long long *regularized, int *retried, int threads) {
    ora_ctx c;
    ora_init(&c, nl);
    size_t *first = malloc(sizeof(size_t) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) {
        size_t k = 0;
        while (k < i && abs(*(int32_t *)(l + (i + k * ld) * c.stride + 8 * nl)) == KIND_ZERO) k++;
        first[i] = k;
    }
    ora_ldl_opts o = {reg_eps != NULL, reg_eps, reg_delta, signs, nsigns, 0, 0, threads};
    int ok;
    if (retried) *retried = 0;
    if (mode == 1) {
        ok = ora_ldl_rec(&c, l, ld, n, first, &o, dinv);
    } else if (mode == 2 && n >= 256) {
        uint8_t *save = malloc(c.stride * n * n);
        for (size_t j = 0; j < n; j++) memcpy(save + j * n * c.stride, l + j * ld * c.stride, n * c.stride);
        ok = ora_ldl_rec(&c, l, ld, n, first, &o, dinv);
        if (!ok) {
            for (size_t j = 0; j < n; j++) memcpy(l + j * ld * c.stride, save + j * n * c.stride, n * c.stride);
            o.regularized = 0;
            if (retried) *retried = 1;
            ok = ora_ldl_dots(&c, l, ld, n, first, &o, dinv);
        }
        free(save);
    } else {
        ok = ora_ldl_dots(&c, l, ld, n, first, &o, dinv);
    }
    if (regularized) *regularized = o.regularized;
    free(first);
    return ok;
}

/* mode as above with recursive != 0 meaning factor_loaded (2). */
int oracle_ldl(int nl, uint8_t *l, size_t ld, size_t n, const int8_t *signs, size_t nsigns,
               const uint8_t *reg_eps, const uint8_t *reg_delta, uint8_t *dinv, int recursive,
               long long *regularized, int threads) {
    return oracle_ldl_mode(nl, l, ld, n, signs, nsigns, reg_eps, reg_delta, dinv,
                           recursive ? 2 : 0, regularized, NULL, threads);
}

/* ------------------------------------------------------------------------
 * Sampled Schur build + arrow factor restatement (sampled/mod.rs:1836-2143,
 * 1973-1988; arrow.rs:985-1019, 1522-1640, 1935-2074, 3378-3404).
 * ---------------------------------------------------------------------- */
typedef struct {
    uint64_t side, samples;
    const void *rinv, *basis, *weights, *coupling;
} ora_leaf;

/* One leaf's sampled Schur block: V = Rinv Q, Gram = V^T V (each entry
 * exact and rounded once), H_pq = RN(RN(RN(w_p w_q) 1/2) RN(K_pq^2 + K_pq^2))
 * (entry, sampled/mod.rs:1973-1988) in the lower triangle of H (K x K),
 * zeros above. */
static void ora_sampled_h(const ora_ctx *c, const ora_leaf *lf, uint8_t *H, int threads) {
    const int nl = c->n;
    size_t st = c->stride;
    size_t h = lf->side, K = lf->samples;
    uint8_t half[8 * 32 + 8];
    memset(half, 0, st);
    ((uint64_t *)half)[nl - 1] = 1ull << 63;
    *(int32_t *)(half + 8 * nl) = KIND_REG;
    *(int32_t *)(half + 8 * nl + 4) = 0;
    uint8_t *V = malloc(st * h * K), *G = malloc(st * (K * (K + 1) / 2));
    ora_gemm(nl, 'N', 'N', h, K, h, lf->rinv, h, lf->basis, h, 0, 0, V, threads);
    ora_gemm(nl, 'T', 'N', K, K, h, V, h, V, h, 1, 1, G, threads);
    for (size_t e = 0; e < K * K; e++) write_zero(H, st, e, nl);
    for (size_t q = 0; q < K; q++)
        for (size_t p = q; p < K; p++) {
            uint8_t t1[8 * 32 + 8], t2[8 * 32 + 8], u1[8 * 32 + 8], u[8 * 32 + 8];
            const uint8_t *kpq = AT(G, p * (p + 1) / 2 + q);
            ora_mul(c, AT((const uint8_t *)lf->weights, p), AT((const uint8_t *)lf->weights, q), t1);
            ora_mul(c, t1, half, t2);
            ora_mul(c, kpq, kpq, u1);
            {
                const uint8_t *xs[2] = {u1, u1}, *ys[2] = {c->one, c->one};
                ora_dot_ptrs(c, 2, xs, ys, u);
            }
            ora_mul(c, t2, u, AT(H, p + q * K));
        }
    free(V);
    free(G);
}

/* The sampled Schur blocks alone: leaf b's H (lower triangle, zeros above)
 * into hout[b]. */
int oracle_sampled_schur(int nl, size_t nleaves, const ora_leaf *leaves, uint8_t **hout, int threads) {
    ora_ctx cc;
    ora_init(&cc, nl);
    for (size_t b = 0; b < nleaves; b++) ora_sampled_h(&cc, &leaves[b], hout[b], threads);
    return 1;
}

int oracle_sampled_arrow(int nl, size_t nleaves, const ora_leaf *leaves, size_t t, uint8_t *border,
                         const int8_t *bsigns, size_t nbsigns, uint8_t *border_dinv, uint8_t **lout,
                         uint8_t **dout, uint8_t **yout, int threads) {
    ora_ctx cc;
    ora_init(&cc, nl);
    const ora_ctx *c = &cc;
    size_t st = c->stride;
    /* Leaves of order 256 and more factor recursively and solve Y by
     * recursive panels in the reference (factor_loaded, arrow.rs:819;
     * forward_panel_rows, arrow.rs:1596-1632); not restated here. */
    for (size_t b = 0; b < nleaves; b++)
        if (leaves[b].samples >= 256) return -1;
    for (size_t b = 0; b < nleaves; b++) {
        const ora_leaf *lf = &leaves[b];
        size_t K = lf->samples;
        uint8_t *H = calloc(K * K, st);
        ora_sampled_h(c, lf, H, threads);
        uint8_t *dinv = malloc(st * K);
        long long reg = 0;
        int8_t plus = 1;
        if (!oracle_ldl(nl, H, K, K, &plus, 1, NULL, NULL, dinv, 0, &reg, threads)) {
            free(H);
            free(dinv);
            return 0;
        }
        /* envelope of the factor's lower triangle (as the reference keeps it) */
        size_t *first = malloc(sizeof(size_t) * K);
        for (size_t i = 0; i < K; i++) {
            size_t k = 0;
            while (k < i && ora_is_zero(c, AT(H, i + k * K))) k++;
            first[i] = k;
        }
        uint8_t *Y = malloc(st * K * t);
        memcpy(Y, lf->coupling, st * K * t);
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : 1)
        for (long long ci = 0; ci < (long long)t; ci++) {
            uint8_t *col = Y + (size_t)ci * K * st;
            const uint8_t **xs = malloc(sizeof(uint8_t *) * (K + 1)), **ys = malloc(sizeof(uint8_t *) * (K + 1));
            uint8_t *negs = malloc(st * (K + 1));
            for (size_t i = 0; i < K; i++) {
                size_t lo = first[i], m = 0;
                for (size_t k = lo; k < i; k++, m++) {
                    memcpy(negs + m * st, AT(H, i + k * K), st);
                    ora_neg(c, negs + m * st);
                    xs[m] = negs + m * st;
                    ys[m] = col + k * st;
                }
                xs[m] = c->one;
                ys[m] = col + i * st;
                m++;
                uint8_t r[8 * 32 + 8];
                ora_dot_ptrs(c, (int)m, xs, ys, r);
                memcpy(col + i * st, r, st);
            }
            free(xs);
            free(ys);
            free(negs);
        }
        free(first);
        uint8_t *Z = malloc(st * K * t), *P = malloc(st * (t * (t + 1) / 2));
        for (size_t e = 0; e < K * t; e++) ora_mul(c, AT(Y, e), AT(dinv, e % K), AT(Z, e));
        ora_gemm(nl, 'T', 'N', t, t, K, Y, K, Z, K, 1, 1, P, threads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
        for (long long jj = 0; jj < (long long)t; jj++) {
            size_t j = (size_t)jj;
            for (size_t i = 0; i <= j; i++) {
                uint8_t r[8 * 32 + 8];
                ora_sub(c, AT(border, i + j * t), AT(P, j * (j + 1) / 2 + i), r);
                memcpy(AT(border, i + j * t), r, st);
            }
        }
        free(Z);
        free(P);
        if (lout && lout[b]) memcpy(lout[b], H, st * K * K);
        if (dout && dout[b]) memcpy(dout[b], dinv, st * K);
        if (yout && yout[b]) memcpy(yout[b], Y, st * K * t);
        free(H);
        free(dinv);
        free(Y);
    }
    for (size_t j = 0; j < t; j++)
        for (size_t i = j + 1; i < t; i++) memcpy(AT(border, i + j * t), AT(border, j + i * t), st);
    long long reg = 0;
    /* factor_border (arrow.rs:3378-3404) through factor_loaded */
    return oracle_ldl_mode(nl, border, t, t, bsigns, nbsigns, NULL, NULL, border_dinv, 2, &reg, NULL, threads);
}

/* ------------------------------------------------------------------------
 * Arrow solves for the sampled arrow factor: ArrowLDLSolver::solve_many
 * (arrow.rs:3793-3980) with the leaves' first_many / second_many
 * (1682-1745) and DenseLeaf's FMA sweeps (forward_many 1155-1164,
 * backward 1165-1183, solve_many 1184-1195), for leaves whose every row
 * couples (coupling_start 0).  x (n x nrhs, ld ldx, n = sum K + t; leaf
 * rows in leaf order, then the t border rows) holds b on entry, x on
 * return.  Per column:
 *   w_l = L_l^-1 b_l          forward: w_i = RN(w_i - L_ik w_k), k ascending
 *   v_l = RN(w_l D_l^-1)
 *   y_j = c_j, then per leaf in order y_j = RN(y_j - RN(sum_i Y_l[i,j] v_l[i]))
 *   y   = L_b^-T RN(D_b^-1 L_b^-1 y)   (border sweeps, backward blocks of 32)
 *   x_l = L_l^-T RN(RN(w_l - RN(Y_l y)) D_l^-1)   (backward blocks of 16)
 * The backward sweep of an order-n factor with block `back` takes, for row
 * i, the blocks [n - back, n), [n - 2 back, n - back), ... down to row
 * i + 1, ascending within a block (DenseLeaf::backward, arrow.rs:1165-1183).
 * ---------------------------------------------------------------------- */
static void ora_forward_fma(const ora_ctx *c, const uint8_t *L, size_t ld, size_t n, uint8_t *x) {
    size_t st = c->stride;
    uint8_t t[8 * 32 + 8];
    for (size_t i = 0; i < n; i++)
        for (size_t k = 0; k < i; k++) {
            ora_fms(c, AT(x, i), AT(L, i + k * ld), AT(x, k), t);
            memcpy(AT(x, i), t, st);
        }
}

static void ora_backward_fma(const ora_ctx *c, const uint8_t *L, size_t ld, size_t n, size_t back, uint8_t *x) {
    size_t st = c->stride;
    uint8_t t[8 * 32 + 8];
    for (size_t ii = n; ii-- > 0;) {
        size_t r1 = n;
        while (r1 > ii + 1) {
            size_t r0 = r1 > back ? r1 - back : 0;
            for (size_t k = (r0 > ii + 1 ? r0 : ii + 1); k < r1; k++) {
                ora_fms(c, AT(x, ii), AT(L, k + ii * ld), AT(x, k), t);
                memcpy(AT(x, ii), t, st);
            }
            r1 = r0;
        }
    }
}

int oracle_sampled_arrow_solve(int nl, size_t nleaves, const uint64_t *ks, size_t t, const uint8_t *l_all,
                               const uint8_t *dinv_all, const uint8_t *y_all, const uint8_t *border,
                               const uint8_t *border_dinv, size_t nrhs, uint8_t *x, size_t ldx, int threads) {
    ora_ctx cc;
    ora_init(&cc, nl);
    const ora_ctx *c = &cc;
    size_t st = c->stride, ksum = 0, kmax = 1;
    for (size_t b = 0; b < nleaves; b++) {
        ksum += ks[b];
        if (ks[b] > kmax) kmax = ks[b];
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (long long cl = 0; cl < (long long)nrhs; cl++) {
        uint8_t *col = x + (size_t)cl * ldx * st;
        uint8_t *w = malloc(st * (ksum ? ksum : 1)), *v = malloc(st * (ksum ? ksum : 1));
        uint8_t *yb = col + ksum * st;
        const uint8_t **xs = malloc(sizeof(uint8_t *) * (kmax + t + 1)),
                      **ys = malloc(sizeof(uint8_t *) * (kmax + t + 1));
        uint8_t t1[8 * 32 + 8], t2[8 * 32 + 8];
        size_t lo = 0, ko = 0;
        /* first: w = L^-1 b, v = RN(w D^-1) per leaf */
        for (size_t b = 0; b < nleaves; b++) {
            size_t K = ks[b];
            memcpy(w + ko * st, col + ko * st, K * st);
            ora_forward_fma(c, l_all + lo * st, K, K, w + ko * st);
            for (size_t i = 0; i < K; i++) ora_mul(c, AT(w, ko + i), AT(dinv_all, ko + i), AT(v, ko + i));
            lo += K * K;
            ko += K;
        }
        /* couple: y_j -= RN(Y_l^T v_l) leaf by leaf */
        for (size_t j = 0; j < t; j++) {
            size_t kk = 0;
            for (size_t b = 0; b < nleaves; b++) {
                size_t K = ks[b];
                for (size_t i = 0; i < K; i++) {
                    xs[i] = AT(y_all, kk * t + i + j * K);
                    ys[i] = AT(v, kk + i);
                }
                ora_dot_ptrs(c, (int)K, xs, ys, t1);
                ora_sub(c, AT(yb, j), t1, t2);
                memcpy(AT(yb, j), t2, st);
                kk += K;
            }
        }
        /* trunk: L_b D_b L_b^T y = y */
        ora_forward_fma(c, border, t, t, yb);
        for (size_t j = 0; j < t; j++) {
            ora_mul(c, AT(yb, j), AT(border_dinv, j), t1);
            memcpy(AT(yb, j), t1, st);
        }
        ora_backward_fma(c, border, t, t, 32, yb);
        /* second: x_l = L^-T RN(RN(w - RN(Y y)) D^-1) */
        lo = 0;
        ko = 0;
        for (size_t b = 0; b < nleaves; b++) {
            size_t K = ks[b];
            for (size_t i = 0; i < K; i++) {
                for (size_t j = 0; j < t; j++) {
                    xs[j] = AT(y_all, ko * t + i + j * K);
                    ys[j] = AT(yb, j);
                }
                ora_dot_ptrs(c, (int)t, xs, ys, t1);
                ora_sub(c, AT(w, ko + i), t1, t2);
                ora_mul(c, t2, AT(dinv_all, ko + i), AT(col, ko + i));
            }
            ora_backward_fma(c, l_all + lo * st, K, K, 16, col + ko * st);
            lo += K * K;
            ko += K;
        }
        free(w);
        free(v);
        free(xs);
        free(ys);
    }
    return 1;
}

/* ------------------------------------------------------------------------
 * Iterative refinement of the sampled arrow KKT solve: DirectLDLKKTSolver::
 * solve_many (direct/solver.rs:470-520) followed by its stationary refine
 * (1046-1129) with the exact-rows residual (ExactRows::residual_negated,
 * 1486-1670): K = [blockdiag(H_l) B; B^T C] with H_l the leaves' sampled
 * Schur blocks (lower triangles, h_all), B_l = coupling (K_l x t each, in
 * b_all leaf after leaf) and C the border block (upper triangle of c, t x t).
 * Per column c:
 *   x = K^-1 b (the arrow solve above); e = RN(b - K x) per row, one exact
 *   dot over (b_i, 1) and (K_ij, -x_j); normb = |b|_inf, norme = |e|_inf
 *   (infinite when e holds a NaN or infinity); ok = finite(norme)
 *   repeat max_iter times, columns in lockstep:
 *     a column stops once norme <= RN(abstol + RN(reltol normb));
 *     dx = K^-1 e; dx = RN(x + dx); e = RN(b - K dx) (exact rows)
 *     non-finite norme: ok = 0, stop; ratio = RN(last / norme);
 *     ratio < stop_ratio stops the column, and keeps x when ratio <= 1;
 *     otherwise x = dx.
 * norms receive the residual norm of the returned x per column;
 * *refinements the correction solves over all columns.
 * ---------------------------------------------------------------------- */
static int ora_cmp_abs(const ora_ctx *c, const uint8_t *a, const uint8_t *b) {
    int ka = abs(*(const int32_t *)(a + 8 * c->n)), kb = abs(*(const int32_t *)(b + 8 * c->n));
    int ca = ka == KIND_ZERO ? 0 : (ka == KIND_REG ? 1 : 2), cb = kb == KIND_ZERO ? 0 : (kb == KIND_REG ? 1 : 2);
    if (ca != cb) return ca < cb ? -1 : 1;
    if (ca != 1) return 0;
    int32_t ea = *(const int32_t *)(a + 8 * c->n + 4), eb = *(const int32_t *)(b + 8 * c->n + 4);
    if (ea != eb) return ea < eb ? -1 : 1;
    for (int i = c->n - 1; i >= 0; i--) {
        uint64_t x = ((const uint64_t *)a)[i], y = ((const uint64_t *)b)[i];
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}
/* sign of a - b, neither NaN (+0 == -0) */
static int ora_cmp(const ora_ctx *c, const uint8_t *a, const uint8_t *b) {
    int sa = ora_is_zero(c, a) ? 0 : (ora_is_neg(c, a) ? -1 : 1);
    int sb = ora_is_zero(c, b) ? 0 : (ora_is_neg(c, b) ? -1 : 1);
    if (sa != sb) return sa < sb ? -1 : 1;
    return sa >= 0 ? ora_cmp_abs(c, a, b) : -ora_cmp_abs(c, a, b);
}
/* |v|_inf of n values, +infinity when one is NaN or infinite */
static void ora_norm_inf(const ora_ctx *c, const uint8_t *v, size_t n, uint8_t *out) {
    size_t st = c->stride;
    write_zero(out, st, 0, c->n);
    for (size_t i = 0; i < n; i++) {
        const uint8_t *x = AT(v, i);
        if (!ora_is_finite(c, x)) {
            memset(out, 0, st);
            *(int32_t *)(out + 8 * c->n) = KIND_INF;
            return;
        }
        if (ora_cmp_abs(c, x, out) > 0) {
            memcpy(out, x, st);
            *(int32_t *)(out + 8 * c->n) = abs(*(int32_t *)(out + 8 * c->n));
        }
    }
}

typedef struct {
    size_t nleaves, t, ksum, n;
    const uint64_t *ks;
    const uint8_t *h_all, *b_all, *cmat;
} ora_kkt;

/* e = RN(b - K x) row by row (x, b, e: n values each) */
static void ora_kkt_residual(const ora_ctx *c, const ora_kkt *k, const uint8_t *b, const uint8_t *x, uint8_t *e,
                             int threads) {
    size_t st = c->stride, n = k->n, t = k->t, ksum = k->ksum;
    uint8_t *neg = malloc(st * n);
    memcpy(neg, x, st * n);
    for (size_t i = 0; i < n; i++) ora_neg(c, AT(neg, i));
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (long long ii = 0; ii < (long long)n; ii++) {
        size_t i = (size_t)ii;
        const uint8_t **xs = malloc(sizeof(uint8_t *) * (n + 1)), **ys = malloc(sizeof(uint8_t *) * (n + 1));
        size_t m = 0;
        xs[m] = AT(b, i);
        ys[m++] = c->one;
        if (i < ksum) {
            size_t l = 0, ko = 0, ho = 0;
            while (i >= ko + k->ks[l]) {
                ko += k->ks[l];
                ho += k->ks[l] * k->ks[l];
                l++;
            }
            size_t K = k->ks[l], r = i - ko;
            const uint8_t *H = k->h_all + ho * st;
            for (size_t q = 0; q < K; q++) {
                xs[m] = r >= q ? AT(H, r + q * K) : AT(H, q + r * K);
                ys[m++] = AT(neg, ko + q);
            }
            const uint8_t *B = k->b_all + ko * t * st;
            for (size_t j = 0; j < t; j++) {
                xs[m] = AT(B, r + j * K);
                ys[m++] = AT(neg, ksum + j);
            }
        } else {
            size_t j = i - ksum, ko = 0;
            for (size_t l = 0; l < k->nleaves; l++) {
                size_t K = k->ks[l];
                const uint8_t *B = k->b_all + ko * t * st;
                for (size_t r = 0; r < K; r++) {
                    xs[m] = AT(B, r + j * K);
                    ys[m++] = AT(neg, ko + r);
                }
                ko += K;
            }
            for (size_t q = 0; q < t; q++) {
                xs[m] = j <= q ? AT(k->cmat, j + q * t) : AT(k->cmat, q + j * t);
                ys[m++] = AT(neg, ksum + q);
            }
        }
        ora_dot_ptrs(c, (int)m, xs, ys, AT(e, i));
        free(xs);
        free(ys);
    }
    free(neg);
}

/* e = RN(b - K x) of one point (the exact-rows residual alone). */
void oracle_sampled_residual(int nl, size_t nleaves, const uint64_t *ks, size_t t, const uint8_t *h_all,
                             const uint8_t *b_all, const uint8_t *cmat, const uint8_t *b, const uint8_t *x, uint8_t *e,
                             int threads) {
    ora_ctx cc;
    ora_init(&cc, nl);
    size_t ksum = 0;
    for (size_t l = 0; l < nleaves; l++) ksum += ks[l];
    ora_kkt k = {nleaves, t, ksum, ksum + t, ks, h_all, b_all, cmat};
    ora_kkt_residual(&cc, &k, b, x, e, threads);
}

int oracle_sampled_refine(int nl, size_t nleaves, const uint64_t *ks, size_t t, const uint8_t *h_all,
                          const uint8_t *b_all, const uint8_t *cmat, const uint8_t *l_all, const uint8_t *dinv_all,
                          const uint8_t *y_all, const uint8_t *border, const uint8_t *border_dinv,
                          const uint8_t *reltol, const uint8_t *abstol, const uint8_t *stop_ratio,
                          unsigned max_iter, size_t nrhs, const uint8_t *b, uint8_t *x, int *ok, uint8_t *norms,
                          long long *refinements, int threads) {
    ora_ctx cc;
    ora_init(&cc, nl);
    const ora_ctx *c = &cc;
    size_t st = c->stride, ksum = 0;
    for (size_t l = 0; l < nleaves; l++) ksum += ks[l];
    ora_kkt k = {nleaves, t, ksum, ksum + t, ks, h_all, b_all, cmat};
    size_t n = k.n;
    *refinements = 0;
    if (n == 0 || nrhs == 0) return 1;
    memcpy(x, b, st * n * nrhs);
    oracle_sampled_arrow_solve(nl, nleaves, ks, t, l_all, dinv_all, y_all, border, border_dinv, nrhs, x, n, threads);
    uint8_t *e = malloc(st * n * nrhs), *dx = malloc(st * n * nrhs);
    uint8_t *normb = malloc(st * nrhs), *norme = malloc(st * nrhs), *last = malloc(st * nrhs);
    uint8_t *tol = malloc(st * nrhs);
    int *active = malloc(sizeof(int) * nrhs);
    for (size_t col = 0; col < nrhs; col++) {
        uint8_t t1[8 * 32 + 8];
        ora_norm_inf(c, b + col * n * st, n, AT(normb, col));
        ora_mul(c, reltol, AT(normb, col), t1);
        const uint8_t *xs[2] = {abstol, t1}, *ys[2] = {c->one, c->one};
        ora_dot_ptrs(c, 2, xs, ys, AT(tol, col));
        ora_kkt_residual(c, &k, b + col * n * st, x + col * n * st, e + col * n * st, threads);
        ora_norm_inf(c, e + col * n * st, n, AT(norme, col));
        memcpy(AT(norms, col), AT(norme, col), st);
        ok[col] = active[col] = ora_is_finite(c, AT(norme, col));
    }
    for (unsigned it = 0; it < max_iter; it++) {
        int any = 0;
        for (size_t col = 0; col < nrhs; col++) {
            if (active[col] && ora_cmp(c, AT(norme, col), AT(tol, col)) <= 0) active[col] = 0;
            any |= active[col];
        }
        if (!any) break;
        memcpy(last, norme, st * nrhs);
        for (size_t col = 0; col < nrhs; col++) {
            if (!active[col]) continue;
            (*refinements)++;
            uint8_t *d = dx + col * n * st;
            memcpy(d, e + col * n * st, st * n);
            oracle_sampled_arrow_solve(nl, nleaves, ks, t, l_all, dinv_all, y_all, border, border_dinv, 1, d, n,
                                       threads);
            for (size_t i = 0; i < n; i++) {
                const uint8_t *xs[2] = {AT(x + col * n * st, i), AT(d, i)}, *ys[2] = {c->one, c->one};
                uint8_t r[8 * 32 + 8];
                ora_dot_ptrs(c, 2, xs, ys, r);
                memcpy(AT(d, i), r, st);
            }
            ora_kkt_residual(c, &k, b + col * n * st, d, e + col * n * st, threads);
            ora_norm_inf(c, e + col * n * st, n, AT(norme, col));
            if (!ora_is_finite(c, AT(norme, col))) {
                ok[col] = active[col] = 0;
                continue;
            }
            /* ratio = RN(last / norme); a zero norm gives +infinity, above every bound */
            int below_stop = 0, not_better = 0;
            if (!ora_is_zero(c, AT(norme, col))) {
                uint8_t ratio[8 * 32 + 8];
                ora_div(c, AT(last, col), AT(norme, col), ratio);
                below_stop = ora_cmp(c, ratio, stop_ratio) < 0;
                not_better = ora_cmp(c, ratio, c->one) <= 0;
            }
            if (below_stop) {
                active[col] = 0;
                if (not_better) continue;
            }
            memcpy(x + col * n * st, d, st * n);
            memcpy(AT(norms, col), AT(norme, col), st);
        }
    }
    free(e);
    free(dx);
    free(normb);
    free(norme);
    free(last);
    free(tol);
    free(active);
    return 1;
}

/* Scalar checks for the tests. */
void oracle_recip(int nl, const uint8_t *d, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_recip(&c, d, out);
}
void oracle_sub(int nl, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_sub(&c, a, b, out);
}
void oracle_mul(int nl, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_mul(&c, a, b, out);
}

int oracle_abi(void) { return 1; }

/* ------------------------------------------------------------------------
 * Exact-dot Cholesky / triangular solve / solve (mpfr.rs:726-872,
 * traits.rs:85-113).
 * ---------------------------------------------------------------------- */
int oracle_potrf(int nl, char uplo, size_t n, uint8_t *a, size_t lda) {
    ora_ctx cc, *c = &cc;
    ora_init(c, nl);
    const int lower = uplo == 'L';
    size_t st = c->stride;
    const uint8_t **xs = malloc(sizeof(void *) * (n ? n : 1)), **ys = malloc(sizeof(void *) * (n ? n : 1));
    uint8_t t[8 * 32 + 8], u[8 * 32 + 8];
#define ATF(i, k) (a + (lower ? (i) + (k) * lda : (k) + (i) * lda) * st)
    for (size_t j = 0; j < n; j++) {
        for (size_t k = 0; k < j; k++) xs[k] = ys[k] = ATF(j, k);
        ora_dot_ptrs(c, (int)j, xs, ys, t);
        uint8_t *ajj = a + (j + j * lda) * st;
        ora_sub(c, ajj, t, u);
        memcpy(ajj, u, st);
        if (!ora_is_finite(c, u) || ora_is_zero(c, u) || ora_is_neg(c, u)) {
            free(xs);
            free(ys);
            return (int)j + 1;
        }
        ora_sqrt(c, u, ajj);
        for (size_t i = j + 1; i < n; i++) {
            for (size_t k = 0; k < j; k++) {
                xs[k] = ATF(i, k);
                ys[k] = ATF(j, k);
            }
            ora_dot_ptrs(c, (int)j, xs, ys, t);
            uint8_t *q = ATF(i, j);
            ora_sub(c, q, t, u);
            ora_div(c, u, ajj, q);
        }
    }
#undef ATF
    free(xs);
    free(ys);
    return 0;
}

void oracle_trsm_lower(int nl, size_t n, const uint8_t *a, uint8_t *b, size_t nrhs, int transpose) {
    ora_ctx cc, *c = &cc;
    ora_init(c, nl);
    size_t st = c->stride;
    const uint8_t **xs = malloc(sizeof(void *) * (n ? n : 1)), **ys = malloc(sizeof(void *) * (n ? n : 1));
    uint8_t t[8 * 32 + 8], u[8 * 32 + 8];
    for (size_t r = 0; r < nrhs; r++) {
        uint8_t *col = b + r * n * st;
        for (size_t step = 0; step < n; step++) {
            size_t i = transpose ? n - 1 - step : step;
            int cnt = 0;
            if (transpose) {
                for (size_t k = i + 1; k < n; k++) {
                    xs[cnt] = a + (k + i * n) * st;
                    ys[cnt++] = col + k * st;
                }
            } else {
                for (size_t k = 0; k < i; k++) {
                    xs[cnt] = a + (i + k * n) * st;
                    ys[cnt++] = col + k * st;
                }
            }
            ora_dot_ptrs(c, cnt, xs, ys, t);
            ora_sub(c, col + i * st, t, u);
            ora_div(c, u, a + (i + i * n) * st, col + i * st);
        }
    }
    free(xs);
    free(ys);
}

void oracle_potrs(int nl, char uplo, size_t n, size_t nrhs, const uint8_t *a, size_t lda, uint8_t *b,
                  size_t ldb) {
    ora_ctx cc, *c = &cc;
    ora_init(c, nl);
    size_t st = c->stride;
    const int lower = uplo == 'L';
    uint8_t v[8 * 32 + 8], w[8 * 32 + 8];
    for (size_t r = 0; r < nrhs; r++) {
        uint8_t *col = b + r * ldb * st;
        for (size_t i = 0; i < n; i++) {
            memcpy(v, col + i * st, st);
            for (size_t k = 0; k < i; k++) {
                const uint8_t *av = a + (lower ? i + k * lda : k + i * lda) * st;
                ora_fms(c, v, av, col + k * st, w);
                memcpy(v, w, st);
            }
            ora_div(c, v, a + (i + i * lda) * st, col + i * st);
        }
        for (size_t ii = n; ii-- > 0;) {
            memcpy(v, col + ii * st, st);
            for (size_t k = ii + 1; k < n; k++) {
                const uint8_t *av = a + (lower ? k + ii * lda : ii + k * lda) * st;
                ora_fms(c, v, av, col + k * st, w);
                memcpy(v, w, st);
            }
            ora_div(c, v, a + (ii + ii * lda) * st, col + ii * st);
        }
    }
}

void oracle_div(int nl, const uint8_t *x, const uint8_t *d, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_div(&c, x, d, out);
}
void oracle_fms(int nl, const uint8_t *v, const uint8_t *a, const uint8_t *b, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_fms(&c, v, a, b, out);
}
void oracle_sqrt(int nl, const uint8_t *a, uint8_t *out) {
    ora_ctx c;
    ora_init(&c, nl);
    ora_sqrt(&c, a, out);
}
