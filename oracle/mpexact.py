"""Exact big-integer oracle for the MpFloat<N> exact-product contract.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path imports this module;
it is used by ``tests/``, by ``tests/golden/make_golden.py`` (which writes the
committed fixtures) and by ``__graft_entry__.smoke()`` as the checker.

Number format (crates/arithmetic/src/lib.rs:333-340, dyadic.rs:1-14): a value
is ``limbs[N]`` (u64, least significant first), ``kind`` (MPFR kind code:
0 NaN, 1 Inf, 2 zero, 3 regular; negative kind = negative sign) and an i32
``exponent``; a regular value is ``M * 2**(exponent - 64N)`` with
``2**(64N-1) <= M < 2**(64N)``.  Zero is all-zero limbs, kind 2, exponent 0
(lib.rs:366-373).

Rounding contract (crates/arithmetic/src/exactdot.rs:1-13,
crates/solver/src/algebra/dense/blas/rns_blas.rs:1-21): every product entry
is the exact dot product rounded once, to nearest with ties to even, at the
destination precision.  ``from_scaled_integer`` follows exactdot.rs:1316-1391.

Everything here is plain Python integers, independent of the C restatement
(oracle/oracle.c) and of the GPU library, so the three can check each other.
"""
from fractions import Fraction
from math import isqrt

NAN_KIND, INF_KIND, ZERO_KIND, REGULAR_KIND = 0, 1, 2, 3
# MPFR's default exponent range (lib.rs:337-339): +-(2**30 - 1).
EMAX = (1 << 30) - 1
EMIN = 1 - (1 << 30)
# Largest admitted per-operand exponent spread (rns_blas.rs:40).
MAX_SPREAD = 4096


class Mp:
    """One MpFloat<N> value as (limbs tuple, kind, exponent)."""

    __slots__ = ("n", "limbs", "kind", "exponent")

    def __init__(self, n, limbs, kind, exponent):
        self.n = n
        self.limbs = tuple(int(x) for x in limbs)
        self.kind = int(kind)
        self.exponent = int(exponent)

    # ---- views -------------------------------------------------------
    @property
    def p(self):
        return 64 * self.n

    def mantissa(self):
        m = 0
        for i, l in enumerate(self.limbs):
            m |= l << (64 * i)
        return m

    def is_zero(self):
        return abs(self.kind) == ZERO_KIND

    def is_regular(self):
        return abs(self.kind) == REGULAR_KIND

    def is_finite(self):
        return abs(self.kind) in (ZERO_KIND, REGULAR_KIND)

    def negative(self):
        return self.kind < 0

    def scaled(self):
        """(signed integer S, scale s) with value = S * 2**s; zero -> (0, 0)."""
        if not self.is_regular():
            if self.is_zero():
                return 0, 0
            raise ValueError("non-finite value has no scaled form")
        m = self.mantissa()
        return (-m if self.negative() else m), self.exponent - self.p

    def to_fraction(self):
        s, e = self.scaled()
        return Fraction(s) * (Fraction(2) ** e)

    def key(self):
        """Bitwise identity with all zeros identified (the reference compares
        ``x == y || both zero``, rns_blas_tests.rs:101-105)."""
        if self.is_zero():
            return ("zero",)
        return (self.limbs, self.kind, self.exponent)

    def __repr__(self):
        return f"Mp(n={self.n}, kind={self.kind}, e={self.exponent}, M={self.mantissa():#x})"


def zero(n):
    return Mp(n, [0] * n, ZERO_KIND, 0)


def from_scaled_integer(negative, mag, scale, n):
    """Round ``(-1)**negative * mag * 2**scale`` once, nearest-even, to 64n
    bits (exactdot.rs:1316-1391).  ``mag`` is a non-negative Python int."""
    if mag < 0:
        raise ValueError("mag must be non-negative")
    if mag == 0:
        return zero(n)
    p = 64 * n
    bits = mag.bit_length()
    shift = bits - p
    if shift > 0:
        m = mag >> shift
        rbit = (mag >> (shift - 1)) & 1
        sticky = (mag & ((1 << (shift - 1)) - 1)) != 0
        if rbit and (sticky or (m & 1)):
            m += 1
            if m >> p:
                m = 1 << (p - 1)
                shift += 1
    else:
        m = mag << (-shift)
    exponent = scale + shift + p
    if exponent > EMAX:
        return Mp(n, [0] * n, -INF_KIND if negative else INF_KIND, 0)
    if exponent < EMIN:
        # MPFR's RNDN underflow: below half the smallest positive -> 0.
        return zero(n)
    limbs = [(m >> (64 * i)) & ((1 << 64) - 1) for i in range(n)]
    return Mp(n, limbs, -REGULAR_KIND if negative else REGULAR_KIND, exponent)


def round_fraction(x, n):
    """Correctly rounded (RNE) 64n-bit value of an exact rational."""
    x = Fraction(x)
    if x == 0:
        return zero(n)
    neg = x < 0
    x = abs(x)
    p = 64 * n
    # find e with 2**(e-1) <= x < 2**e
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** e <= x:
        e += 1
    while Fraction(2) ** (e - 1) > x:
        e -= 1
    scaled = x * (Fraction(2) ** (p - e))  # in [2**(p-1), 2**p)
    q, r = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * r
    if twice > scaled.denominator or (twice == scaled.denominator and (q & 1)):
        q += 1
    if q >> p:
        q >>= 1
        e += 1
    return from_scaled_integer(neg, q, e - p, n)


def from_int(v, n):
    return round_fraction(Fraction(v), n)


def exact_dot(pairs, n):
    """Exact sum of products rounded once.  ``pairs`` yields (Mp, Mp)."""
    terms = []
    for a, b in pairs:
        if a.is_zero() or b.is_zero():
            continue
        sa, ea = a.scaled()
        sb, eb = b.scaled()
        terms.append((sa * sb, ea + eb))
    if not terms:
        return zero(n)
    lo = min(e for _, e in terms)
    total = sum(s << (e - lo) for s, e in terms)
    return from_scaled_integer(total < 0, abs(total), lo, n)


def _exponent_range(vals):
    """(min, max) of ``exponent - P`` over non-zero entries; None if any
    entry is non-finite (rns_blas.rs:620-632)."""
    lo, hi = None, None
    for v in vals:
        if not v.is_finite():
            return "nonfinite"
        if v.is_zero():
            continue
        x = v.exponent - v.p
        lo = x if lo is None else min(lo, x)
        hi = x if hi is None else max(hi, x)
    return lo, hi


def gemm_exact(ta, tb, m, n, k, a, lda, b, ldb, upper_only=False, packed=None):
    """The exact product contract of rns_blas.rs:2134-2417.

    ``a`` and ``b`` are flat column-major lists of Mp.  Returns None on a
    decline (non-finite operand entry, or an operand exponent spread above
    MAX_SPREAD), else the output list: full m*n column-major (ldc = m), or
    the column-packed upper triangle when ``packed`` (square, upper_only).
    Upper-only, unpacked outputs return the full list with the strictly lower
    entries left as None (unwritten)."""
    ta = ta.upper() if isinstance(ta, str) else chr(ta).upper()
    tb = tb.upper() if isinstance(tb, str) else chr(tb).upper()
    nl = (a[0].n if a else b[0].n) if (a or b) else 1
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    stored_a = [a[i + j * lda] for j in range(ac) for i in range(ar)] if ar and ac else []
    stored_b = [b[i + j * ldb] for j in range(bc) for i in range(br)] if br and bc else []
    ra, rb = _exponent_range(stored_a), _exponent_range(stored_b)
    if ra == "nonfinite" or rb == "nonfinite":
        return None
    for r in (ra, rb):
        if r[0] is not None and r[1] - r[0] > MAX_SPREAD:
            return None
    if packed is None:
        packed = upper_only and m == n

    def at(i, p):
        return a[i + p * lda] if ta == "N" else a[p + i * lda]

    def bt(p, j):
        return b[p + j * ldb] if tb == "N" else b[j + p * ldb]

    if packed:
        out = []
        for j in range(n):
            for i in range(j + 1):
                out.append(exact_dot(((at(i, p), bt(p, j)) for p in range(k)), nl))
        return out
    out = [None] * (m * n)
    for j in range(n):
        for i in range(m):
            if upper_only and i > j:
                continue
            out[i + j * m] = exact_dot(((at(i, p), bt(p, j)) for p in range(k)), nl)
    return out


# ---------------------------------------------------------------------------
# Scalar correctly rounded operations (used by the LDL^T restatement).
# ---------------------------------------------------------------------------

def mul(x, y):
    if x.is_zero() or y.is_zero():
        return zero(x.n)
    sx, ex = x.scaled()
    sy, ey = y.scaled()
    s = sx * sy
    return from_scaled_integer(s < 0, abs(s), ex + ey, x.n)


def add(x, y):
    if x.is_zero():
        return y
    if y.is_zero():
        return x
    sx, ex = x.scaled()
    sy, ey = y.scaled()
    lo = min(ex, ey)
    s = (sx << (ex - lo)) + (sy << (ey - lo))
    return from_scaled_integer(s < 0, abs(s), lo, x.n)


def neg(x):
    if x.is_zero():
        return x
    return Mp(x.n, x.limbs, -x.kind, x.exponent)


def sub(x, y):
    return add(x, neg(y))


def div(x, y):
    return round_fraction(x.to_fraction() / y.to_fraction(), x.n)


def sqrt(x):
    """RN(sqrt(x)) for x >= 0 (MPFR sqrt, round to nearest): the integer
    square root of the scaled mantissa with at least 2p + 4 bits, the
    remainder as sticky, one rounding."""
    if x.is_zero():
        return zero(x.n)
    s, e = x.scaled()
    if s < 0:
        raise ValueError("sqrt of a negative value")
    if e & 1:
        s <<= 1
        e -= 1
    p = 64 * x.n
    k = max(0, (2 * p + 4 - s.bit_length() + 1) // 2)
    r = s << (2 * k)
    q = isqrt(r)
    sticky = 1 if q * q != r else 0
    return from_scaled_integer(False, (q << 1) | sticky, e // 2 - k - 1, x.n)


def fms(v, a, b):
    """RN(v - a b) with one rounding (MPFR mul_add of -a, b, v)."""
    return exact_dot([(v, from_int(1, v.n)), (neg(a), b)], v.n)


# ---------------------------------------------------------------------------
# Exact-dot Cholesky, triangular solve and solve (mpfr.rs:726-872,
# traits.rs:85-113): every subtracted sum is one exact dot rounded once.
# ---------------------------------------------------------------------------

def potrf(uplo, n, a, lda):
    """XpotrfScalar::xpotrf (mpfr.rs:726-789) on a list of Mp, in place.
    Returns info: 0, or the failing pivot column + 1."""
    nl = a[0].n if a else 1
    lower = uplo == "L"

    def at(i, k):
        return i + k * lda if lower else k + i * lda

    for j in range(n):
        d = sub(a[j + j * lda], exact_dot([(a[at(j, k)], a[at(j, k)]) for k in range(j)], nl))
        a[j + j * lda] = d
        if not d.is_finite() or d.is_zero() or d.negative():
            return j + 1
        d = sqrt(d)
        a[j + j * lda] = d
        for i in range(j + 1, n):
            solved = exact_dot([(a[at(i, k)], a[at(j, k)]) for k in range(j)], nl)
            q = at(i, j)
            a[q] = div(sub(a[q], solved), d)
    return 0


def trsm_lower(n, a, b, transpose):
    """XpotrsScalar::xtrsm_lower (traits.rs:85-113): L x = b (or L^T x = b)
    per column of b (a: n x n column-major lower), in place."""
    if n == 0:
        return
    nl = a[0].n
    for c0 in range(0, len(b), n):
        col = b[c0:c0 + n]
        for step in range(n):
            i = n - 1 - step if transpose else step
            if transpose:
                solved = exact_dot([(a[r + i * n], col[r]) for r in range(i + 1, n)], nl)
            else:
                solved = exact_dot([(a[i + k * n], col[k]) for k in range(i)], nl)
            col[i] = div(sub(col[i], solved), a[i + i * n])
        b[c0:c0 + n] = col


def potrs(uplo, n, nrhs, a, lda, b, ldb):
    """XpotrsScalar::xpotrs (mpfr.rs:790-872): forward then backward FMA
    chains per right-hand side, each step one rounding, then a division."""
    lower = uplo == "L"
    for r in range(nrhs):
        col = [b[i + r * ldb] for i in range(n)]
        for i in range(n):
            v = col[i]
            for k in range(i):
                av = a[i + k * lda] if lower else a[k + i * lda]
                v = fms(v, av, col[k])
            col[i] = div(v, a[i + i * lda])
        for i in reversed(range(n)):
            v = col[i]
            for k in range(i + 1, n):
                av = a[k + i * lda] if lower else a[i + k * lda]
                v = fms(v, av, col[k])
            col[i] = div(v, a[i + i * lda])
        for i in range(n):
            b[i + r * ldb] = col[i]


# ---------------------------------------------------------------------------
# Test-data generator of the reference's residue-BLAS tests
# (crates/solver/src/algebra/dense/blas/rns_blas_tests.rs:4-28).
# ---------------------------------------------------------------------------

class Lcg:
    MUL = 6364136223846793005
    INC = 1442695040888963407
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next(self):
        """Binary64 value in [0, 1), exactly as the reference computes it."""
        self.state = (self.state * self.MUL + self.INC) & self.MASK
        return float(self.state >> 11) / float(1 << 53)


def lcg_random(rng, length, spread, n):
    """rns_blas_tests.rs:16-28: 5% zeros, else (2u-1)/(v+0.5) rounded at 64n
    bits, scaled by 2**shift with shift uniform in [-spread, spread]."""
    out = []
    for _ in range(length):
        # Python floats are IEEE binary64, so every comparison, sum and
        # product below rounds exactly as the Rust f64 code does.
        if rng.next() < 0.05:
            out.append(zero(n))
            continue
        # from_f64 of a binary64 value is exact at >= 64 bits.
        num = round_fraction(Fraction(rng.next() * 2.0 - 1.0), n)
        den = round_fraction(Fraction(rng.next() + 0.5), n)
        shift = int(rng.next() * float(2 * spread + 1)) - spread
        q = div(num, den)
        if q.is_zero():
            out.append(q)
            continue
        out.append(Mp(n, q.limbs, q.kind, q.exponent + shift))
    return out


# ---------------------------------------------------------------------------
# Dense LDL^T restatement (crates/solver/src/solver/kkt/ldl/arrow.rs):
# ldl_dots 156-253, ldl_rec 255-415, solve_unit_lower_t 417-490,
# DenseLeaf::pivot 618-634.  Matrices are dicts/lists indexed [i + j*ld].
# ---------------------------------------------------------------------------
REC_BASE = 128


def _pivot(d, sign, reg, stats):
    if not d.is_finite():
        return None
    if reg is not None:
        eps, delta = reg
        ds = d if sign > 0 else neg(d)
        if ds.to_fraction() < eps.to_fraction():
            d = delta if sign > 0 else neg(delta)
            stats["regularized"] += 1
    if d.is_zero():
        return None
    return d


def ldl_dots(l, ld, n, first, sign, reg, stats, dinv, off=0):
    """In place on the lower triangle of l[off + i + j*ld]; dinv[i]."""
    nl = l[off].n
    one = from_int(1, nl)
    r = {}
    w = {}
    for j in range(n):
        fj = first[j]
        dot = exact_dot(((r[(j, q)], w[(j, q)]) for q in range(fj, j)), nl)
        a_jj = l[off + j + j * ld]
        dj = sub(a_jj, dot)
        if dj.is_zero() and j > fj:
            exact = exact_dot([(r[(j, q)], w[(j, q)]) for q in range(fj, j)] + [(a_jj, neg(one))], nl)
            dj = neg(exact)
        dj = _pivot(dj, sign(j), reg, stats)
        if dj is None:
            return False
        dinv[j] = div(one, dj)
        for i in range(j + 1, n):
            if first[i] > j:
                v = zero(nl)
            else:
                lo = max(first[i], first[j])
                dot = exact_dot(((r[(i, q)], w[(j, q)]) for q in range(lo, j)), nl)
                v = mul(sub(l[off + i + j * ld], dot), dinv[j])
            r[(i, j)] = v
            w[(i, j)] = mul(v, dj)
    for j in range(n):
        l[off + j + j * ld] = one
        for i in range(j + 1, n):
            l[off + i + j * ld] = r[(i, j)]
    return True


def exact_product_nt(m, n, k, a, aoff, lda, b, boff, ldb, nl, upper_packed=False):
    """op = A * B^T (A m x k at a[aoff + i + p*lda], B n x k)."""
    if upper_packed:
        out = []
        for j in range(n):
            for i in range(j + 1):
                out.append(exact_dot(((a[aoff + i + p * lda], b[boff + j + p * ldb]) for p in range(k)), nl))
        return out
    out = [None] * (m * n)
    for j in range(n):
        for i in range(m):
            out[i + j * m] = exact_dot(((a[aoff + i + p * lda], b[boff + j + p * ldb]) for p in range(k)), nl)
    return out


def solve_unit_lower_t(m, n, L, loff, ldl, B, boff, ldb, low, nl):
    if m == 0 or n == 0:
        return
    if n <= REC_BASE:
        for i in range(m):
            x = [B[boff + i + j * ldb] for j in range(n)]
            lo = low[i]
            for j in range(lo, n):
                dot = exact_dot(((x[q], L[loff + j + q * ldl]) for q in range(lo, j)), nl)
                x[j] = sub(x[j], dot)
            for j in range(n):
                B[boff + i + j * ldb] = x[j]
        return
    n1, n2 = n // 2, n - n // 2
    solve_unit_lower_t(m, n1, L, loff, ldl, B, boff, ldb, [min(f, n1) for f in low], nl)
    p = exact_product_nt(m, n2, n1, B, boff, ldb, L, loff + n1, ldl, nl)
    for j in range(n2):
        for i in range(m):
            idx = boff + i + (n1 + j) * ldb
            B[idx] = sub(B[idx], p[i + j * m])
    solve_unit_lower_t(m, n2, L, loff + n1 + n1 * ldl, ldl, B, boff + n1 * ldb, ldb,
                       [max(f - n1, 0) for f in low], nl)


def ldl_rec(l, ld, n, first, sign, reg, stats, dinv, off=0, doff=0):
    nl = l[off].n
    if n <= REC_BASE:
        d = {}
        ok = ldl_dots(l, ld, n, first, sign, reg, stats, d, off)
        for k, v in d.items():
            dinv[doff + k] = v
        return ok
    low, cut = n, None
    for c in range(n - 1, 0, -1):
        low = min(low, first[c])
        if low >= c and (cut is None or abs(c - n // 2) < abs(cut - n // 2)):
            cut = c
    if cut is not None:
        c = cut
        ok1 = ldl_rec(l, ld, c, first[:c], sign, reg, stats, dinv, off, doff)
        ok2 = ldl_rec(l, ld, n - c, [f - c for f in first[c:]], lambda k: sign(c + k), reg, stats,
                      dinv, off + c + c * ld, doff + c)
        return ok1 and ok2
    n1, n2 = n // 2, n - n // 2
    if not ldl_rec(l, ld, n1, first[:n1], sign, reg, stats, dinv, off, doff):
        return False
    l11 = [l[off + i + j * ld] for j in range(n1) for i in range(n1)]
    solve_unit_lower_t(n2, n1, l11, 0, n1, l, off + n1, ld, [min(f, n1) for f in first[n1:]], nl)
    w = [l[off + n1 + i + j * ld] for j in range(n1) for i in range(n2)]
    for j in range(n1):
        for i in range(n2):
            idx = off + n1 + i + j * ld
            l[idx] = mul(l[idx], dinv[doff + j])
    packed = exact_product_nt(n2, n2, n1, w, 0, n2, l, off + n1, ld, nl, upper_packed=True)
    for a in range(n2):
        for b in range(a, n2):
            idx = off + n1 + b + (n1 + a) * ld
            l[idx] = sub(l[idx], packed[b * (b + 1) // 2 + a])
    return ldl_rec(l, ld, n2, [max(f - n1, 0) for f in first[n1:]], lambda k: sign(n1 + k), reg,
                   stats, dinv, off + n1 + n1 * ld, doff + n1)


REC_FACTOR_MIN = 256


def factor_loaded(l, ld, n, sign, reg, stats, dinv):
    """DenseLeaf::factor_loaded for wide types (arrow.rs:808-850): from order
    REC_FACTOR_MIN (arrow.rs:62) ldl_rec; when it fails, the input restored
    and factored by ldl_dots, the failed attempt's regularization count
    dropped (arrow.rs:820-846).  Below that order ldl_dots.  The envelope is
    taken once, from the input.  stats["retried"] records the retry."""
    first = envelope(l, ld, n)
    stats.setdefault("retried", False)
    if n >= REC_FACTOR_MIN:
        saved = list(l)
        count = {"regularized": 0}
        if ldl_rec(l, ld, n, first, sign, reg, count, dinv):
            stats["regularized"] += count["regularized"]
            return True
        l[:] = saved
        stats["retried"] = True
    d = {}
    ok = ldl_dots(l, ld, n, first, sign, reg, stats, d)
    for k, v in d.items():
        dinv[k] = v
    return ok


def envelope(l, ld, n):
    first = []
    for i in range(n):
        k = 0
        while k < i and l[i + k * ld].is_zero():
            k += 1
        first.append(k)
    return first


def reference_leaf_matrix(n, nl):
    """The input of the reference's recursive_leaf_factor test
    (arrow.rs:4276-4300): signs, envelope with a cut and ragged rows, a tiny
    pivot regularized by (1e-20, 1)."""
    signs = [-1 if k % 5 == 4 else 1 for k in range(n)]

    def first_of(i):
        if i < 110:
            return 0
        if i < 230:
            return 110
        return 110 + i % 13

    a = [zero(nl)] * (n * n)
    for j in range(n):
        for i in range(j, n):
            if j < first_of(i):
                continue
            if i == j:
                v = float(signs[i]) * (1e-40 if i % 37 == 3 else float(n + i % 7))
            else:
                v = ((i * 29 + j * 13) % 17) / 17.0 - 0.5
            a[i + j * n] = round_fraction(Fraction(v), nl)
    reg = (round_fraction(Fraction(1e-20), nl), from_int(1, nl))
    return a, signs, reg


# ---------------------------------------------------------------------------
# Exact multilinear forms (crates/solver/src/algebra/dense/blas/rns_blas.rs:
# diag_congruence 2997-3195, congruence 3563-3790, svec_quadratic 4032-4173,
# symmetric_bilinear 4175-4276; the XgemmScalar hooks at blas/traits.rs:
# 259-296).  Every output is the exact value of the form rounded once.  Each
# operand is carried as exact integers over its own common scale, the way
# the reference's residue encoding sees it, so the sums below are exact.
# ---------------------------------------------------------------------------

def _integers(vals):
    """(ints, scale) with vals[i] == ints[i] * 2**scale, or None where the
    reference's exponent-range gate declines (a non-finite entry or a spread
    above MAX_SPREAD, rns_blas.rs:620-632).  An all-zero list gives scale
    None."""
    r = _exponent_range(vals)
    if r == "nonfinite":
        return None
    lo, hi = r
    if lo is None:
        return [0] * len(vals), None
    if hi - lo > MAX_SPREAD:
        return None
    ints = []
    for v in vals:
        if v.is_zero():
            ints.append(0)
        else:
            s, e = v.scaled()
            ints.append(s << (e - lo))
    return ints, lo


def _rounded(total, scale, n):
    if total == 0 or scale is None:
        return zero(n)
    return from_scaled_integer(total < 0, abs(total), scale, n)


def diag_congruence_exact(m, k, a, d, upper_only=True):
    """rns_blas.rs:2997-3195: (a^T diag(d) a)_ij, a k x m column-major
    (ld k), d k values; entries i > j are None when upper_only."""
    nl = (a[0] if a else d[0]).n if (a or d) else 1
    ia, id_ = _integers(a[: k * m]), _integers(d[:k])
    if ia is None or id_ is None:
        return None
    (ia, sa), (idv, sd) = ia, id_
    scale = None if sa is None or sd is None else 2 * sa + sd
    out = [None] * (m * m)
    for j in range(m):
        for i in range(j + 1 if upper_only else m):
            total = sum(ia[t + i * k] * idv[t] * ia[t + j * k] for t in range(k))
            out[i + j * m] = _rounded(total, scale, nl)
    return out


def congruence_exact(ta, m, k, a, lda, x, ldx, upper_only=False):
    """rns_blas.rs:3563-3790: op(a) x op(a)^T (ldc = m), x k x k (ld ldx),
    op(a) m x k; entries i > j are None when upper_only."""
    ta = ta.upper()
    rows, cols = (m, k) if ta == "N" else (k, m)
    av = [a[i + j * lda] for j in range(cols) for i in range(rows)]
    xv = [x[i + j * ldx] for j in range(k) for i in range(k)]
    nl = (av + xv)[0].n if (av or xv) else 1
    ia, ix = _integers(av), _integers(xv)
    if ia is None or ix is None:
        return None
    (ia, sa), (ixv, sx) = ia, ix
    scale = None if sa is None or sx is None else 2 * sa + sx

    def opa(i, p):  # op(a)[i, p] over the view (rows x cols, ld rows)
        return ia[i + p * rows] if ta == "N" else ia[p + i * rows]

    # t[p][j] = (x op(a)^T)[p, j], exact
    t = [[sum(ixv[p + q * k] * opa(j, q) for q in range(k)) for j in range(m)] for p in range(k)]
    out = [None] * (m * m)
    for j in range(m):
        for i in range(j + 1 if upper_only else m):
            out[i + j * m] = _rounded(sum(opa(i, p) * t[p][j] for p in range(k)), scale, nl)
    return out


def svec_quadratic_exact(h, kmax, q, abs_q, x, s):
    """rns_blas.rs:4032-4173: v[k] = sum_{i <= j} c_ij x_t q_ik q_jk with
    t = j(j+1)/2 + i, c_ii = 1, c_ij = s; q h x kmax (ld h), |q| when abs_q.
    None where the reference declines: the operand gates above, and a weight
    s with sigma = exponent - P above 0 or below -(P + 64)
    (rns_blas.rs:4083-4085)."""
    nl = s.n
    tri = h * (h + 1) // 2
    qv = [abs_mp(v) if abs_q else v for v in q[: h * kmax]]
    iq, ix = _integers(qv), _integers(x[:tri])
    if iq is None or ix is None:
        return None
    (iq, sq), (ixv, sx) = iq, ix
    if sq is None or sx is None:
        return [zero(nl)] * kmax
    if s.is_zero():
        s_int, sigma = 0, 0
    else:
        if not s.is_regular():
            return None
        s_int, sigma = s.scaled()
        if sigma > 0 or -sigma > s.p + 64:
            return None
    out = []
    for c in range(kmax):
        total = 0
        for j in range(h):
            for i in range(j + 1):
                w = (1 << -sigma) if i == j else s_int
                total += w * ixv[j * (j + 1) // 2 + i] * iq[i + c * h] * iq[j + c * h]
        out.append(_rounded(total, sigma + sx + 2 * sq, nl))
    return out


def symmetric_bilinear_exact(h, columns, q, x, pairs):
    """rns_blas.rs:4175-4276: q_a^T X q_b for (a, b) in pairs, X symmetric
    given as its column-packed upper triangle (no svec factors), q h x
    columns (ld h).  None for h = 0 or a pair index >= columns, as there."""
    if h == 0 or any(a >= columns or b >= columns for a, b in pairs):
        return None
    nl = q[0].n
    tri = h * (h + 1) // 2
    iq, ix = _integers(q[: h * columns]), _integers(x[:tri])
    if iq is None or ix is None:
        return None
    (iq, sq), (ixv, sx) = iq, ix
    if sq is None or sx is None:
        return [zero(nl)] * len(pairs)

    def xe(i, j):
        i, j = min(i, j), max(i, j)
        return ixv[j * (j + 1) // 2 + i]

    out = []
    for a, b in pairs:
        total = 0
        for i in range(h):
            tb = sum(xe(i, j) * iq[j + b * h] for j in range(h))
            total += iq[i + a * h] * tb
        out.append(_rounded(total, sx + 2 * sq, nl))
    return out


def abs_mp(v):
    if v.kind < 0:
        return Mp(v.n, v.limbs, -v.kind, v.exponent)
    return v


def scalar_code(v):
    """The reference's scalar tests in axpby and gemm (F == 0, F == 1,
    F == -1, blas/mpfr.rs:61-80): 0, 1, 2, else 3."""
    if v.is_zero():
        return 0
    if v.is_regular():
        f = v.to_fraction()
        if f == 1:
            return 1
        if f == -1:
            return 2
    return 3


def axpby(alpha, x, beta, y):
    """blas/mpfr.rs:61-80: p = 0, x, -x or RN(alpha x); then p, RN(p + y),
    RN(p - y) or RN(p + RN(beta y))."""
    a, b = scalar_code(alpha), scalar_code(beta)
    p = zero(x.n) if a == 0 else x if a == 1 else neg(x) if a == 2 else mul(alpha, x)
    if b == 0:
        return p
    if b == 1:
        return add(p, y)
    if b == 2:
        return sub(p, y)
    return add(p, mul(beta, y))
