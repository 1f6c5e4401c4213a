"""Multiple-precision dense LDL^T: restatements pinned, product bit-exact.

* The Python restatement (oracle/mpexact.py, plain integers) and the C
  oracle agree bitwise on the reference's own recursive_leaf_factor input
  (arrow.rs:4276-4353: n = 300 at 128 bits, signs, an envelope with a cut
  and ragged rows, tiny pivots regularized by (1e-20, 1)).
* The reference's invariants on that input hold: pivots were regularized,
  the recursive factor differs bitwise from ldl_dots but agrees within
  1e-25 relative, envelope zeros stay exact zeros.
* The library's host-product path (the same recursion with host exact dots
  in place of GPU products) is bitwise the C oracle on CPU.
"""
import numpy as np
import pytest

import mpexact as M
import sdpx_gpu
from helpers import from_mp, oracle_ldl
from sdpx_gpu import mp


def reference_input(n=300, nl=2):
    a, signs, reg = M.reference_leaf_matrix(n, nl)
    return from_mp(a, nl), signs, (from_mp([reg[0]], nl), from_mp([reg[1]], nl)), a, reg


def test_python_and_c_restatements_agree_on_reference_input():
    n, nl = 300, 2
    A, signs, reg, a, reg_mp = reference_input(n, nl)
    ok, L, dinv, count = oracle_ldl(A, n, signs, reg, recursive=True)
    assert ok and count > 0
    l = list(a)
    d = {}
    stats = {"regularized": 0}
    first = M.envelope(l, n, n)
    assert M.ldl_rec(l, n, n, first, lambda k: signs[k], reg_mp, stats, d)
    assert stats["regularized"] == count
    assert np.all(mp.same_bits(from_mp(l, nl), L))
    assert np.all(mp.same_bits(from_mp([d[k] for k in range(n)], nl), dinv))


def test_reference_invariants_recursive_vs_dots():
    n, nl = 300, 2
    A, signs, reg, _, _ = reference_input(n, nl)
    ok1, Lr, dr, c1 = oracle_ldl(A, n, signs, reg, recursive=True)
    ok2, Ld, dd, c2 = oracle_ldl(A, n, signs, reg, recursive=False)
    assert ok1 and ok2 and c1 == c2 and c1 > 0
    assert not np.all(mp.same_bits(Lr, Ld))  # assert_ne!(serial.l, dots.l)
    xr, xd = mp.to_float64(Lr).reshape(n, n).T, mp.to_float64(Ld).reshape(n, n).T
    fr, fd = mp.to_float64(dr), mp.to_float64(dd)
    # |x - y| <= 1e-25 max(|y|, 1): float64 views carry 53 bits, so compare
    # the exact difference through the oracle's subtraction instead
    diff = M.Fraction(0)
    worst = 0.0
    for idx in np.flatnonzero(~mp.same_bits(Lr, Ld)):
        x = M.Mp(nl, Lr[idx]["limbs"], Lr[idx]["kind"], Lr[idx]["exponent"]).to_fraction()
        y = M.Mp(nl, Ld[idx]["limbs"], Ld[idx]["kind"], Ld[idx]["exponent"]).to_fraction()
        worst = max(worst, float(abs(x - y)) / max(float(abs(y)), 1.0))
    assert worst <= 1e-25
    del diff, xr, xd, fr, fd
    # envelope: row i of L is zero before first(i)
    first = [0 if i < 110 else (110 if i < 230 else 110 + i % 13) for i in range(n)]
    kinds = Lr["kind"].reshape(n, n).T  # [i, j]
    for i in range(n):
        for j in range(min(first[i], i)):
            assert abs(kinds[i, j]) == 2


def test_host_product_path_is_the_oracle():
    n, nl = 300, 2
    A, signs, reg, _, _ = reference_input(n, nl)
    ok, L, dinv, count = oracle_ldl(A, n, signs, reg, recursive=True)
    okp, Lp, dp = sdpx_gpu.mp_ldl(A, n, signs, reg, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert okp and np.all(mp.same_bits(Lp, L)) and np.all(mp.same_bits(dp, dinv))
    assert sdpx_gpu.mp_ldl_stats()["regularized"] == count
    okd, Ld, dd = sdpx_gpu.mp_ldl(A, n, signs, reg,
                                  flags=sdpx_gpu.LDL_HOST_PRODUCTS | sdpx_gpu.LDL_DOTS)
    ok2, L2, d2, _ = oracle_ldl(A, n, signs, reg, recursive=False)
    assert okd and np.all(mp.same_bits(Ld, L2)) and np.all(mp.same_bits(dd, d2))


@pytest.mark.parametrize("nl,n", [(4, 520), (8, 300)])
def test_host_product_path_random_spd(nl, n):
    rng = np.random.default_rng(n)
    g = rng.standard_normal((n, n))
    s = g @ g.T / n + 0.05 * np.eye(n)
    A = mp.from_float64(s.T.reshape(-1), nl)
    ok, L, dinv, _ = oracle_ldl(A, n)
    okp, Lp, dp = sdpx_gpu.mp_ldl(A, n, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert ok and okp
    assert np.all(mp.same_bits(Lp, L)) and np.all(mp.same_bits(dp, dinv))


def test_failed_pivot_reported():
    nl, n = 4, 260
    s = -np.eye(n)
    s[0, 0] = 0.0
    A = mp.from_float64(s.T.reshape(-1), nl)
    okp, _, _ = sdpx_gpu.mp_ldl(A, n, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    ok, _, _, _ = oracle_ldl(A, n)
    assert okp is False and ok is False
