"""Iterative refinement of the sampled arrow KKT solve (CPU): the oracle's
exact-rows residual is RN(b - K x) row by row for the KKT assembled from
the leaves' sampled Schur blocks, couplings and border block
(direct/solver.rs:1486-1670); its refinement loop (refine, 1046-1129)
reaches the exact solution of K x = b (mpmath at 4096 bits) to the working
precision from a factored solve that misses it by far more; and its
stopping rules (max_iter, the tolerance, the stop ratio) act as stated."""
from fractions import Fraction

import mpmath
import numpy as np

import mpexact as M
import pins
from helpers import (oracle_sampled_arrow, oracle_sampled_arrow_solve, oracle_sampled_refine,
                     oracle_sampled_residual, oracle_sampled_schur)
from sdpx_gpu import mp, workloads

NL = 4
P = 64 * NL


def small():
    return workloads.sampled_problem(NL, nblocks=3, hmin=4, hmax=7, t=9, seed=13)


def kkt_fractions(prob, h_all, mirror_h=True):
    """K = [blockdiag(H) B; B^T C] as a dense list of Fraction rows from the
    records (H: lower triangles; C: the border's upper triangle)."""
    ks = [b["k"] for b in prob["blocks"]]
    t = prob["t"]
    ksum = sum(ks)
    n = ksum + t
    K = [[Fraction(0)] * n for _ in range(n)]
    hv = pins.dec(h_all)
    ho = ko = 0
    for blk, k in zip(prob["blocks"], ks):
        for q in range(k):
            for p in range(q, k):
                K[ko + p][ko + q] = hv[ho + p + q * k]
                if mirror_h:
                    K[ko + q][ko + p] = hv[ho + p + q * k]
        bv = pins.dec(blk["b"])
        for j in range(t):
            for r in range(k):
                K[ko + r][ksum + j] = K[ksum + j][ko + r] = bv[r + j * k]
        ho += k * k
        ko += k
    cv = pins.dec(prob["border"])
    for q in range(t):
        for p in range(q + 1):
            K[ksum + p][ksum + q] = K[ksum + q][ksum + p] = cv[p + q * t]
    return K


def test_oracle_residual_is_the_exact_row_residual():
    prob = small()
    h_all = oracle_sampled_schur(prob, NL)
    K = kkt_fractions(prob, h_all)
    n = len(K)
    rng = np.random.default_rng(5)
    b = mp.from_float64(rng.standard_normal(n), NL)
    x = mp.from_float64(rng.standard_normal(n) * 1e3, NL)
    e = oracle_sampled_residual(prob, h_all, b, x)
    bv, xv = pins.dec(b), pins.dec(x)
    want = [pins.rn(bv[i] - sum(K[i][j] * xv[j] for j in range(n)), P) for i in range(n)]
    assert pins.dec(e) == want
    # a KKT whose leaf blocks keep only their stored lower triangles (no
    # mirror) gives other rows: the comparison sees the difference
    Kbad = kkt_fractions(prob, h_all, mirror_h=False)
    bad = [pins.rn(bv[i] - sum(Kbad[i][j] * xv[j] for j in range(n)), P) for i in range(n)]
    assert bad != want


def default_tolerance():
    """eps^(3/4) as the reference forms its MPFR default (settings.rs:389-396):
    root = RN(sqrt(eps)), RN(root RN(sqrt(root))), eps = 2^(1-P)."""
    eps = M.round_fraction(Fraction(2) ** (1 - P), NL)
    root = M.sqrt(eps)
    return M.mul(root, M.sqrt(root))


def rec_of(mpval):
    from helpers import from_mp

    return from_mp([mpval], NL)


def factor(prob):
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, NL)
    assert ok
    return border, bdinv, outs


def test_refinement_reaches_the_exact_solution():
    prob = small()
    h_all = oracle_sampled_schur(prob, NL)
    border, bdinv, outs = factor(prob)
    K = kkt_fractions(prob, h_all)
    n, nrhs = len(K), 2
    rng = np.random.default_rng(2)
    rhs = mp.from_float64(rng.standard_normal(n * nrhs), NL)
    x0 = oracle_sampled_arrow_solve(outs, prob["t"], border, bdinv, rhs, nrhs)
    tiny = pins.enc([Fraction(1, 2 ** 400)], NL)
    x, ok, norms, count = oracle_sampled_refine(prob, h_all, outs, border, bdinv, rhs, nrhs, tiny, tiny, 5.0, 10)
    assert ok == [True, True] and count >= nrhs
    mpmath.mp.prec = 4096
    A = mpmath.matrix([[mpmath.mpf(v.numerator) / v.denominator for v in row] for row in K])
    bv, g0, g = pins.dec(rhs), pins.dec(x0), pins.dec(x)
    for c in range(nrhs):
        xs = mpmath.lu_solve(A, mpmath.matrix([mpmath.mpf(v.numerator) / v.denominator
                                               for v in bv[c * n:(c + 1) * n]]))
        scale = max(abs(xs[i]) for i in range(n))
        err0 = max(abs(mpmath.mpf(g0[c * n + i].numerator) / g0[c * n + i].denominator - xs[i]) for i in range(n))
        err = max(abs(mpmath.mpf(g[c * n + i].numerator) / g[c * n + i].denominator - xs[i]) for i in range(n))
        # the factored solve loses some 60 bits to the leaves' conditioning;
        # refinement against the exact residual recovers the working precision
        assert err0 / scale > mpmath.mpf(2) ** -240, float(err0 / scale)
        assert err / scale < mpmath.mpf(2) ** -250, float(err / scale)
        # the returned norm is the returned point's exact-rows residual
        e = oracle_sampled_residual(prob, h_all, rhs[c * n:(c + 1) * n], x[c * n:(c + 1) * n])
        assert pins.dec(norms[c:c + 1])[0] == max(abs(v) for v in pins.dec(e))


def test_refinement_stopping_rules():
    prob = small()
    h_all = oracle_sampled_schur(prob, NL)
    border, bdinv, outs = factor(prob)
    n = sum(b["k"] for b in prob["blocks"]) + prob["t"]
    nrhs = 3
    rng = np.random.default_rng(9)
    rhs = mp.from_float64(rng.standard_normal(n * nrhs), NL)
    x0 = oracle_sampled_arrow_solve(outs, prob["t"], border, bdinv, rhs, nrhs)
    tiny = pins.enc([Fraction(1, 2 ** 400)], NL)
    # no iterations: the factored solve itself
    x, ok, _, count = oracle_sampled_refine(prob, h_all, outs, border, bdinv, rhs, nrhs, tiny, tiny, 5.0, 0)
    assert count == 0 and np.all(mp.same_bits(x, x0))
    # a tolerance every residual meets: no correction
    x, ok, _, count = oracle_sampled_refine(prob, h_all, outs, border, bdinv, rhs, nrhs, 1.0, 1.0, 5.0, 10)
    assert count == 0 and np.all(mp.same_bits(x, x0))
    # the reference's default tolerances at 256 bits (eps^(3/4) absolute and
    # relative): the factored solve's residual misses them (the leaves'
    # conditioning), and the refined columns end below abstol + reltol |b|
    tol = default_tolerance()
    x, ok, norms, count = oracle_sampled_refine(prob, h_all, outs, border, bdinv, rhs, nrhs, rec_of(tol),
                                                rec_of(tol), 5.0, 10)
    assert ok == [True] * nrhs and count >= nrhs
    bv = pins.dec(rhs)
    for c in range(nrhs):
        normb = max(abs(v) for v in bv[c * n:(c + 1) * n])
        bound = M.add(tol, M.mul(tol, M.round_fraction(normb, NL))).to_fraction()
        assert pins.dec(norms[c:c + 1])[0] <= bound
    # a stop ratio no step reaches: exactly one correction per column
    huge = pins.enc([Fraction(2) ** 1000], NL)
    x, ok, norms, count = oracle_sampled_refine(prob, h_all, outs, border, bdinv, rhs, nrhs, tiny, tiny, huge, 10)
    assert count == nrhs and ok == [True] * nrhs
    assert not np.all(mp.same_bits(x, x0))
