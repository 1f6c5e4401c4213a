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

int oracle_abi(void) { return 1; }
