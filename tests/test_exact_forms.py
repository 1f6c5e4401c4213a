"""The oracle's exact multilinear forms (oracle/mpexact.py, restating
rns_blas.rs:2997-3195, 3563-3790, 4032-4173, 4175-4276) pinned by what the
mathematics fixes: a value that rounding once and rounding twice separate,
the off-diagonal weight of svec_quadratic, integer closed forms, identities
that reduce one form to another, and the exact product of the C oracle on
operands small enough that the intermediate products are representable."""
from fractions import Fraction

import pytest

import mpexact
import pins
from exact_forms_cases import lcg_values, rounding_once_case, sqrt2_of, weight_case
from helpers import from_mp, oracle_gemm, to_mp


def vals_of(fracs, nl):
    return to_mp(pins.enc(fracs, nl))


def values(out):
    return [None if v is None else v.to_fraction() for v in out]


@pytest.mark.parametrize("nl", [2, 4, 8])
def test_every_form_rounds_the_exact_value_once(nl):
    p = 64 * nl
    a, d, once, _ = rounding_once_case(p)
    one = Fraction(1)
    # diag_congruence: a = (a, 1) (k = 2, m = 1), d = (d, -1)
    out = mpexact.diag_congruence_exact(1, 2, vals_of([a, one], nl), vals_of([d, -one], nl))
    assert values(out) == [once]
    # congruence: op(a) = (a, 1) (m = 1, k = 2), x = diag(d, -1)
    for ta, lda in (("N", 1), ("T", 2)):
        out = mpexact.congruence_exact(ta, 1, 2, vals_of([a, one], nl), lda, vals_of([d, 0, 0, -one], nl), 2)
        assert values(out) == [once]
    # svec_quadratic: q = (a, 1), x = (d, 0, -1)
    s = vals_of([sqrt2_of(nl)], nl)[0]
    out = mpexact.svec_quadratic_exact(2, 1, vals_of([a, one], nl), False, vals_of([d, 0, -one], nl), s)
    assert values(out) == [once]
    # symmetric_bilinear: q = (a, 1), X = diag(d, -1), pair (0, 0)
    out = mpexact.symmetric_bilinear_exact(2, 1, vals_of([a, one], nl), vals_of([d, 0, -one], nl), [(0, 0)])
    assert values(out) == [once]


def test_rounding_once_case_separates_a_rounding_per_product():
    """The planted mistake: the diagonal scaling rounded before the sum (as a
    diag(d) a product stored at working precision would be)."""
    nl = 2
    p = 64 * nl
    a, d, once, twice = rounding_once_case(p)
    ad = pins.rn(a * d, p)
    planted = mpexact.gemm_exact("T", "N", 1, 1, 2, vals_of([a, 1], nl), 2, vals_of([ad, -1], nl), 2)
    assert values(planted) == [twice] != [once]


@pytest.mark.parametrize("nl", [2, 4])
def test_svec_weights(nl):
    q, x, s, expected = weight_case(nl)
    out = mpexact.svec_quadratic_exact(2, 1, vals_of(q, nl), False, vals_of(x, nl), vals_of([s], nl)[0])
    assert values(out) == [expected]
    # the same form read from |q| with q = (-1, 1): |q| = (1, 1)
    out = mpexact.svec_quadratic_exact(2, 1, vals_of([-1, 1], nl), True, vals_of(x, nl), vals_of([s], nl)[0])
    assert values(out) == [expected]
    # and without |.|: 1 - s^2
    out = mpexact.svec_quadratic_exact(2, 1, vals_of([-1, 1], nl), False, vals_of(x, nl), vals_of([s], nl)[0])
    assert values(out) == [pins.rn(1 - s * s, 64 * nl)]


def test_integer_closed_forms():
    nl, m, k, h = 2, 5, 7, 6
    ones = lambda n: vals_of([1] * n, nl)
    # a^T diag(1..k) a with a all ones: k (k + 1) / 2 everywhere, upper only
    out = mpexact.diag_congruence_exact(m, k, ones(k * m), vals_of(list(range(1, k + 1)), nl))
    for j in range(m):
        for i in range(m):
            assert (out[i + j * m] is None) == (i > j)
            if i <= j:
                assert out[i + j * m].to_fraction() == k * (k + 1) // 2
    # op(a) x op(a)^T with a = ones (m x k), x = ones: k^2
    out = mpexact.congruence_exact("N", m, k, ones(m * k), m, ones(k * k), k)
    assert values(out) == [k * k] * (m * m)
    # q^T X q with q = (1..h), X = ones: (h (h + 1) / 2)^2, by both forms
    q = vals_of(list(range(1, h + 1)), nl)
    xs = ones(h * (h + 1) // 2)
    expect = (h * (h + 1) // 2) ** 2
    assert values(mpexact.symmetric_bilinear_exact(h, 1, q, xs, [(0, 0)])) == [expect]
    two = vals_of([2], nl)[0]
    assert values(mpexact.svec_quadratic_exact(h, 1, q, False, xs, two)) == [expect]


@pytest.mark.parametrize("nl", [2, 4])
def test_identities_between_forms(nl):
    """congruence with x = diag(d) and op(a) = a^T is diag_congruence;
    congruence with a = I returns x; svec_quadratic with weight 2 on an
    unscaled packed X is symmetric_bilinear's (k, k) pairs; diag_congruence
    with a = I is diag(d)."""
    m, k, h, kmax = 6, 9, 7, 5
    a = lcg_values(k * m, 20, 11, nl)
    d = lcg_values(k, 20, 12, nl)
    xd = [mpexact.zero(nl)] * (k * k)
    for t in range(k):
        xd[t + t * k] = d[t]
    got = mpexact.congruence_exact("T", m, k, a, k, xd, k, upper_only=True)
    want = mpexact.diag_congruence_exact(m, k, a, d)
    assert [v and v.key() for v in got] == [v and v.key() for v in want]
    x = lcg_values(k * k, 20, 13, nl)
    eye = [mpexact.from_int(1 if i == j else 0, nl) for j in range(k) for i in range(k)]
    got = mpexact.congruence_exact("N", k, k, eye, k, x, k)
    assert [v.key() for v in got] == [v.key() for v in x]
    got = mpexact.diag_congruence_exact(k, k, eye, d, upper_only=False)
    assert values(got) == [d[i].to_fraction() if i == j else 0 for j in range(k) for i in range(k)]
    q = lcg_values(h * kmax, 20, 14, nl)
    xs = lcg_values(h * (h + 1) // 2, 20, 15, nl)
    two = mpexact.from_int(2, nl)
    quad = mpexact.svec_quadratic_exact(h, kmax, q, False, xs, two)
    bil = mpexact.symmetric_bilinear_exact(h, kmax, q, xs, [(c, c) for c in range(kmax)])
    assert [v.key() for v in quad] == [v.key() for v in bil]


def small_values(count, seed, nl, bits=12):
    """Integers of at most `bits` bits times powers of two: every product of
    three of them is exact at 64 nl bits."""
    rng = mpexact.Lcg(seed)
    out = []
    for _ in range(count):
        v = int(rng.next() * (1 << bits)) - (1 << (bits - 1))
        e = int(rng.next() * 9) - 4
        out.append(Fraction(v) * Fraction(2) ** e)
    return out


def test_forms_match_exact_products_of_representable_intermediates():
    """With small operands diag(d) a and x op(a)^T are exact at P bits, so
    each form is one exact product of the C oracle: a transposed operand, a
    wrong index or a dropped term in the forms fails this."""
    nl, m, k, h, cols = 2, 7, 9, 8, 5
    a, d = small_values(k * m, 1, nl), small_values(k, 2, nl)
    da = [d[t] * a[t + i * k] for i in range(m) for t in range(k)]
    want = oracle_gemm("T", "N", m, m, k, pins.enc(a, nl), k, pins.enc(da, nl), k)
    got = mpexact.diag_congruence_exact(m, k, vals_of(a, nl), vals_of(d, nl), upper_only=False)
    assert values(got) == pins.dec(want)
    # congruence, op(a) = a (m x k) and a^T (a stored k x m)
    x = small_values(k * k, 3, nl)
    for ta in "NT":
        an = small_values(m * k, 4, nl)
        rows = m if ta == "N" else k
        opa = lambda i, p: an[i + p * m] if ta == "N" else an[p + i * k]
        t = [sum(x[p + q * k] * opa(j, q) for q in range(k)) for j in range(m) for p in range(k)]
        opa_m = [opa(i, p) for p in range(k) for i in range(m)]
        want = oracle_gemm("N", "N", m, m, k, pins.enc(opa_m, nl), m, pins.enc(t, nl), k)
        got = mpexact.congruence_exact(ta, m, k, vals_of(an, nl), rows, vals_of(x, nl), k)
        assert values(got) == pins.dec(want), ta
    # symmetric_bilinear: q_a^T (X q_b)
    q = small_values(h * cols, 5, nl)
    xs = small_values(h * (h + 1) // 2, 6, nl)
    X = lambda i, j: xs[max(i, j) * (max(i, j) + 1) // 2 + min(i, j)]
    pairs = [(0, 1), (3, 3), (4, 2), (1, 0)]
    got = mpexact.symmetric_bilinear_exact(h, cols, vals_of(q, nl), vals_of(xs, nl), pairs)
    for (ia, ib), v in zip(pairs, got):
        xq = [sum(X(i, j) * q[j + ib * h] for j in range(h)) for i in range(h)]
        want = oracle_gemm("T", "N", 1, 1, h, pins.enc(q[ia * h:(ia + 1) * h], nl), h, pins.enc(xq, nl), h)
        assert v.to_fraction() == pins.dec(want)[0]


def test_declines_follow_the_reference():
    nl = 2
    q = vals_of([1, 1], nl)
    xs = vals_of([1, 1, 1], nl)
    assert mpexact.symmetric_bilinear_exact(0, 1, [], [], [(0, 0)]) is None
    assert mpexact.symmetric_bilinear_exact(2, 1, q, xs, [(0, 1)]) is None
    big = vals_of([Fraction(2) ** 130], nl)[0]  # sigma = exponent - P > 0
    assert mpexact.svec_quadratic_exact(2, 1, q, False, xs, big) is None
    spread = vals_of([1, Fraction(2) ** 5000], nl)
    assert mpexact.diag_congruence_exact(1, 2, spread, vals_of([1, 1], nl)) is None
    zeros = [mpexact.zero(nl)] * 4
    assert values(mpexact.congruence_exact("N", 2, 2, zeros, 2, vals_of([1, 2, 3, 4], nl), 2)) == [0] * 4


def axpby_case(p):
    """alpha = 1 + u, x = 1 + 3u, beta = -1, y = 1 (u = 2^(1-p)): p =
    RN(alpha x) = 1 + 4u and p - y = 4u, where the exact alpha x - y =
    4u + 3u^2 rounds once to 4u + 4u^2.  Returns (alpha, x, beta, y, want,
    fused)."""
    u = Fraction(1, 2 ** (p - 1))
    alpha, x, beta, y = 1 + u, 1 + 3 * u, Fraction(-1), Fraction(1)
    want, fused = 4 * u, pins.rn(alpha * x - y, p)
    assert fused == 4 * u + 4 * u * u != want
    return alpha, x, beta, y, want, fused


@pytest.mark.parametrize("nl", [2, 4])
def test_axpby_rounds_the_scaled_product_before_the_sum(nl):
    alpha, x, beta, y, want, _ = axpby_case(64 * nl)
    a, xv, b, yv = vals_of([alpha, x, beta, y], nl)
    assert mpexact.axpby(a, xv, b, yv).to_fraction() == want
    # the special scalars: 0, 1, -1 for alpha; 0, 1, -1 for beta; other beta
    one, mone, z = vals_of([1, -1, 0], nl)
    assert mpexact.axpby(z, xv, z, yv).is_zero()
    assert mpexact.axpby(one, xv, z, yv).to_fraction() == x
    assert mpexact.axpby(mone, xv, one, yv).to_fraction() == pins.rn(-x + y, 64 * nl)
    assert mpexact.axpby(one, xv, mone, yv).to_fraction() == pins.rn(x - y, 64 * nl)
    b3 = vals_of([Fraction(3, 4)], nl)[0]
    assert mpexact.axpby(one, xv, b3, xv).to_fraction() == pins.rn(x + pins.rn(Fraction(3, 4) * x, 64 * nl), 64 * nl)
