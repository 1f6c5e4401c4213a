"""Inputs and known answers of the exact multilinear forms (diag_congruence,
congruence, svec_quadratic, symmetric_bilinear), shared by the CPU pins of
the oracle and the GPU parity tests."""
from fractions import Fraction
from math import isqrt

import mpexact
import pins
from helpers import from_mp


def rounding_once_case(p):
    """One form value a d a - 1 with a = d = 1 + u, u = 2^(1-p).  The exact
    value 3u + 3u^2 + u^3 lies just above the midpoint 3u + 3u^2 of the
    p-bit neighbours 3u + 2u^2 and 3u + 4u^2, so rounding once gives
    3u + 4u^2; rounding a d first (to 1 + 2u) gives 3u + 2u^2 instead.
    Returns (a, d, expected, twice)."""
    u = Fraction(1, 2 ** (p - 1))
    a = d = 1 + u
    exact = a * d * a - 1
    assert exact == 3 * u + 3 * u * u + u ** 3
    once = 3 * u + 4 * u * u
    twice = 3 * u + 2 * u * u
    assert pins.rn(exact, p) == once and pins.rn(pins.rn(a * d, p) * a - 1, p) == twice
    return a, d, once, twice


def sqrt2_of(nl):
    """RN(sqrt 2) at 64 nl bits, from the integer square root."""
    p = 64 * nl
    r = isqrt(2 << (2 * p))  # floor(sqrt2 * 2^p)
    # nearest: compare (r + 1/2)^2 with 2^(2p+1)
    if (2 * r + 1) ** 2 < (8 << (2 * p)):
        r += 1
    v = Fraction(r, 2 ** p)
    return pins.rn(v, p)


def weight_case(nl):
    """svec_quadratic with h = 2, one column q = (1, 1), svec x = (1, s, 0):
    v = 1 * x_00 + s * x_01 = 1 + s^2, rounded once; c_00 = s instead gives
    s + s^2 and c_01 = 2s gives 1 + 2 s^2.  Returns (q, x, s, expected)."""
    p = 64 * nl
    s = sqrt2_of(nl)
    q = [Fraction(1), Fraction(1)]
    x = [Fraction(1), s, Fraction(0)]
    expected = pins.rn(1 + s * s, p)
    assert expected != pins.rn(s + s * s, p) and expected != pins.rn(1 + 2 * s * s, p)
    return q, x, s, expected


def lcg_values(count, spread, seed, nl):
    return mpexact.lcg_random(mpexact.Lcg(seed), count, spread, nl)


def random_forms(nl, spread, seed, m=20, k=33, h=16, kmax=18, cols=11, npairs=25):
    """Random operands (the reference's test distribution,
    rns_blas_tests.rs:16-28) of every form, as oracle values."""
    rng = mpexact.Lcg(seed)
    vals = lambda n: mpexact.lcg_random(rng, n, spread, nl)
    pairs = [((7 * i + 3) % cols, (5 * i + 1) % cols) for i in range(npairs)]
    return {
        "diag": (m, k, vals(k * m), vals(k)),
        "cong_n": (m, k, vals(m * k), vals(k * k)),
        "cong_t": (m, k, vals(k * m), vals(k * k)),
        "quad": (h, kmax, vals(h * kmax), vals(h * (h + 1) // 2)),
        "bil": (h, cols, vals(h * cols), vals(h * (h + 1) // 2), pairs),
    }


def records(vals, nl):
    return from_mp(vals, nl)
