"""The oracle's LDL^T and sampled arrow factor pinned by known answers fixed
by mathematics (tests/pins.py): the C oracle, the Python restatement and
the library's host-panel path must all give the closed-form bits.  A
planted mistake in the restatement (a dropped term of the dot, a naive
quotient in place of the rounding sequence) fails these cases."""
from fractions import Fraction

import numpy as np
import pytest

import mpexact as M
import pins
import sdpx_gpu
from helpers import oracle_ldl, oracle_sampled_arrow, to_mp, from_mp
from sdpx_gpu import mp


def check_factor(l_rec, dinv_rec, n, L_expect, dinv_expect):
    """l_rec's lower triangle is the unit lower L_expect ([i][j], j < i),
    dinv_rec is dinv_expect, bit for bit (as values: records are canonical)."""
    got = pins.lower_of(l_rec, n)
    for i in range(n):
        assert got[i][i] == 1
        for j in range(i):
            assert got[i][j] == L_expect[i][j], (i, j)
    assert pins.dec(dinv_rec) == list(dinv_expect)


def records_of(A, nl):
    return pins.enc(pins.col_major(A), nl)


@pytest.mark.parametrize("n", [3, 130, 300])
def test_exact_ldl_known_answer(n):
    nl = 2
    A, L, D, signs = pins.exact_ldl_case(n)
    rec = records_of(A, nl)
    dinv_expect = [1 / d for d in D]
    ok, l, dinv, _ = oracle_ldl(rec, n, signs)  # factor_loaded: recursive from 256
    assert ok
    check_factor(l, dinv, n, L, dinv_expect)
    okd, ld, dd, _ = oracle_ldl(rec, n, signs, recursive=False)
    assert okd
    check_factor(ld, dd, n, L, dinv_expect)
    okh, lh, dh = sdpx_gpu.mp_ldl(rec, n, signs, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert okh
    check_factor(lh, dh, n, L, dinv_expect)
    if n <= 130:
        lp = to_mp(rec)
        d = {}
        assert M.factor_loaded(lp, n, n, lambda k: signs[k], None, {"regularized": 0}, d)
        check_factor(from_mp(lp, nl), from_mp([d[k] for k in range(n)], nl), n, L, dinv_expect)


@pytest.mark.parametrize("nl", [2, 4, 8])
def test_rounding_order_known_answer(nl):
    p = 64 * nl
    A, r, dinv = pins.rounding_order_case(p)
    rec = records_of(A, nl)
    L = [[r[i][j] if j < i else None for j in range(3)] for i in range(3)]
    ok, l, d, _ = oracle_ldl(rec, 3, recursive=False)
    assert ok
    check_factor(l, d, 3, L, dinv)
    okh, lh, dh = sdpx_gpu.mp_ldl(rec, 3, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert okh
    check_factor(lh, dh, 3, L, dinv)
    lp = to_mp(rec)
    dd = {}
    assert M.ldl_dots(lp, 3, 3, [0, 0, 0], lambda k: 1, None, {"regularized": 0}, dd)
    check_factor(from_mp(lp, nl), from_mp([dd[k] for k in range(3)], nl), 3, L, dinv)


@pytest.mark.parametrize("nl", [2, 4])
def test_zero_pivot_exact_retry(nl):
    p = 64 * nl
    A, x, dinv1 = pins.zero_pivot_case(p)
    rec = records_of(A, nl)
    L = [[None, None], [x, None]]
    ok, l, d, _ = oracle_ldl(rec, 2, recursive=False)
    assert ok
    check_factor(l, d, 2, L, [Fraction(1), dinv1])
    okh, lh, dh = sdpx_gpu.mp_ldl(rec, 2, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert okh
    check_factor(lh, dh, 2, L, [Fraction(1), dinv1])


def test_recursive_factor_retried_by_exact_dots():
    nl, n = 2, 256
    A, x, dinv = pins.rec_retry_case(64 * nl)
    rec = records_of(A, nl)
    L = [[Fraction(0)] * n for _ in range(n)]
    L[128][0] = x
    # the raw recursive factor fails on pivot 128
    raw_ok, _, _, _ = oracle_ldl(rec, n, mode="rec")
    assert not raw_ok
    ok, l, d, _ = oracle_ldl(rec, n)  # factor_loaded
    assert ok
    check_factor(l, d, n, L, dinv)
    okh, lh, dh = sdpx_gpu.mp_ldl(rec, n, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    st = sdpx_gpu.mp_ldl_stats()
    assert okh and st["retried"] == 1
    check_factor(lh, dh, n, L, dinv)


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("perturbed", [False, True])
def test_analytic_sampled_schur(nl, perturbed):
    p = 64 * nl
    case = pins.analytic_sampled_case(p, p // 4 if perturbed else None)
    H = case["H"]
    if not perturbed:
        assert H == [[16, 0], [0, 36]]  # sampled/tests.rs:424-426
    else:
        delta = Fraction(1, 2 ** (p // 4))
        cross = -(2 * delta + delta * delta)
        assert H[0][1] == -6 * cross * cross  # sampled/tests.rs:428-433, exact here
    prob = pins.sampled_problem_of(case, nl)
    expect_l = [[None, None], [case["L10"], None]]
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl)
    assert ok
    l, d, y = outs[0]
    check_factor(l, d, 2, expect_l, case["dinv"])
    assert pins.dec(y) == case["y"]
    # the border's lower triangle holds its unit factor, D^-1 = RN(1/s)
    assert pins.dec(border) == [1] and pins.dec(bdinv) == [case["border_dinv"]]
    okh, bh, dh, fh = sdpx_gpu.sampled_arrow(prob, nl, flags=sdpx_gpu.LDL_HOST_PRODUCTS,
                                            want_factors=True)
    assert okh
    lh, dhh, yh = fh[0]
    check_factor(lh, dhh, 2, expect_l, case["dinv"])
    assert pins.dec(yh) == case["y"]
    assert pins.dec(bh) == [1] and pins.dec(dh) == [case["border_dinv"]]


def test_planted_mistakes_fail_the_pins(monkeypatch):
    """Mistakes planted in the Python restatement (on the CPU) fail the
    known answers: a dot that drops its last term, and products rounded
    toward zero instead of to nearest-even."""
    nl = 2
    A, L, D, signs = pins.exact_ldl_case(12)
    rec = records_of(A, nl)
    exact_dot = M.exact_dot

    def dropped(pairs, n):
        pairs = list(pairs)
        return exact_dot(pairs[:-1], n)

    monkeypatch.setattr(M, "exact_dot", dropped)
    lp = to_mp(rec)
    d = {}
    M.ldl_dots(lp, 12, 12, M.envelope(lp, 12, 12), lambda k: signs[k], None, {"regularized": 0}, d)
    got = pins.lower_of(from_mp(lp, nl), 12)
    assert any(got[i][j] != L[i][j] for i in range(12) for j in range(i))
    monkeypatch.setattr(M, "exact_dot", exact_dot)

    A3, r, dinv = pins.rounding_order_case(64 * nl)
    rec3 = records_of(A3, nl)

    def mul_toward_zero(x, y):
        if x.is_zero() or y.is_zero():
            return M.zero(x.n)
        v = x.to_fraction() * y.to_fraction()
        t = M.round_fraction(v, x.n)
        if abs(t.to_fraction()) > abs(v):  # rounded away: one unit back toward zero
            ulp = Fraction(2) ** (t.exponent - 64 * x.n)
            t = M.round_fraction(t.to_fraction() - (ulp if v > 0 else -ulp), x.n)
        return t

    monkeypatch.setattr(M, "mul", mul_toward_zero)
    lp = to_mp(rec3)
    dd = {}
    M.ldl_dots(lp, 3, 3, [0, 0, 0], lambda k: 1, None, {"regularized": 0}, dd)
    got = pins.lower_of(from_mp(lp, nl), 3)
    assert got[1][0] != r[1][0] or got[2][1] != r[2][1] or got[2][0] != r[2][0]


def test_sampled_border_retried_by_exact_dots():
    nl = 2
    prob, x, dinv = pins.retry_border_problem(nl)
    t = prob["t"]
    L = [[Fraction(0)] * t for _ in range(t)]
    L[128][0] = x
    ok, border, bdinv, _ = oracle_sampled_arrow(prob, nl)
    assert ok
    check_factor(border, bdinv, t, L, dinv)
    okh, bh, dh, _ = sdpx_gpu.sampled_arrow(prob, nl, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert okh
    check_factor(bh, dh, t, L, dinv)
