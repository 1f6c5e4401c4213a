"""The known answers of tests/pins.py on the GPU paths: the device LDL^T
(sdpx_gpu_mp_ldl through its device panels) and the device sampled arrow
pipeline give the closed-form bits, and report that no step declined to
the host panels."""
from fractions import Fraction

import pytest

import pins
import sdpx_gpu
from test_pins import check_factor, records_of

pytestmark = pytest.mark.gpu


def device_ldl(rec, n, signs=None):
    ok, l, d = sdpx_gpu.mp_ldl(rec, n, signs)
    st = sdpx_gpu.mp_ldl_stats()
    assert st["declined"] == 0, "device panels declined"
    return ok, l, d, st


@pytest.mark.parametrize("n", [3, 130, 300])
def test_exact_ldl_known_answer_gpu(gpu, n):
    nl = 2
    print("exact ldl", n, flush=True)
    A, L, D, signs = pins.exact_ldl_case(n)
    ok, l, d, _ = device_ldl(records_of(A, nl), n, signs)
    assert ok
    check_factor(l, d, n, L, [1 / v for v in D])


@pytest.mark.parametrize("nl", [2, 4, 8])
def test_rounding_order_known_answer_gpu(gpu, nl):
    A, r, dinv = pins.rounding_order_case(64 * nl)
    L = [[r[i][j] if j < i else None for j in range(3)] for i in range(3)]
    ok, l, d, _ = device_ldl(records_of(A, nl), 3)
    assert ok
    check_factor(l, d, 3, L, dinv)


@pytest.mark.parametrize("nl", [2, 4])
def test_zero_pivot_exact_retry_gpu(gpu, nl):
    A, x, dinv1 = pins.zero_pivot_case(64 * nl)
    ok, l, d, _ = device_ldl(records_of(A, nl), 2)
    assert ok
    check_factor(l, d, 2, [[None, None], [x, None]], [Fraction(1), dinv1])


def test_recursive_factor_retried_by_exact_dots_gpu(gpu):
    nl, n = 2, 256
    print("rec retry", n, flush=True)
    A, x, dinv = pins.rec_retry_case(64 * nl)
    ok, l, d, st = device_ldl(records_of(A, nl), n)
    assert ok and st["retried"] == 1
    L = [[Fraction(0)] * n for _ in range(n)]
    L[128][0] = x
    check_factor(l, d, n, L, dinv)


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("perturbed", [False, True])
def test_analytic_sampled_schur_gpu(gpu, nl, perturbed):
    p = 64 * nl
    case = pins.analytic_sampled_case(p, p // 4 if perturbed else None)
    prob = pins.sampled_problem_of(case, nl)
    ok, border, bdinv, fac = sdpx_gpu.sampled_arrow(prob, nl, want_factors=True)
    st = sdpx_gpu.sampled_stats()
    assert ok and st["declined"] == 0 and st["gpu_products"] > 0
    l, d, y = fac[0]
    check_factor(l, d, 2, [[None, None], [case["L10"], None]], case["dinv"])
    assert pins.dec(y) == case["y"]
    assert pins.dec(border) == [1] and pins.dec(bdinv) == [case["border_dinv"]]


def test_sampled_border_retried_by_exact_dots_gpu(gpu):
    nl = 2
    print("sampled border retry", flush=True)
    prob, x, dinv = pins.retry_border_problem(nl)
    t = prob["t"]
    ok, border, bdinv, _ = sdpx_gpu.sampled_arrow(prob, nl)
    st = sdpx_gpu.sampled_stats()
    assert ok and st["declined"] == 0 and st["retried"] == 1
    L = [[Fraction(0)] * t for _ in range(t)]
    L[128][0] = x
    check_factor(border, bdinv, t, L, dinv)
