"""Sampled Schur build + arrow factor (one IPM iteration of the sampled
bootstrap-style SDP): the library's host-product path is bitwise the C
oracle's restatement on a reduced instance (CPU), and the analytic known
answers of the sampled Schur entry hold (sampled/tests.rs:407-434)."""
import numpy as np

import mpexact as M
import sdpx_gpu
from helpers import oracle_sampled_arrow
from sdpx_gpu import mp, workloads


def test_analytic_schur_entries():
    # identity R, basis [1, 1; 1, -1], weights (2, -3): Gram = Q^T Q =
    # diag(2, 2); entry(p, q) = w_p w_q K_pq^2 -> 16, 36, 0 exactly
    # (sampled/tests.rs:407-434)
    nl = 4
    q = np.array([[1.0, 1.0], [1.0, -1.0]])  # columns are basis vectors
    gram = q.T @ q
    w = np.array([2.0, -3.0])
    ent = lambda p, r: M.mul(M.mul(M.mul(M.from_int(int(w[p]), nl), M.from_int(int(w[r]), nl)),
                                   M.round_fraction(M.Fraction(1, 2), nl)),
                             M.add(M.from_int(int(gram[p, r] ** 2), nl), M.from_int(int(gram[p, r] ** 2), nl)))
    assert ent(0, 0).to_fraction() == 16 and ent(1, 1).to_fraction() == 36
    assert ent(0, 1).is_zero()


def small_problem(nl, seed=3):
    return workloads.sampled_problem(nl, nblocks=4, hmin=6, hmax=14, t=150, seed=seed)


def test_host_path_matches_oracle():
    nl = 4
    prob = small_problem(nl)
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl)
    okp, bp, dp, fp = sdpx_gpu.sampled_arrow(prob, nl, flags=sdpx_gpu.LDL_HOST_PRODUCTS,
                                             want_factors=True)
    assert ok and okp
    for (l, d, y), (lp, dpp, yp) in zip(outs, fp):
        assert np.all(mp.same_bits(lp, l))
        assert np.all(mp.same_bits(dpp, d))
        assert np.all(mp.same_bits(yp, y))
    t = prob["t"]
    lower = np.tril_indices(t)
    idx = lower[0] + lower[1] * t
    assert np.all(mp.same_bits(bp[idx], border[idx]))
    assert np.all(mp.same_bits(dp, bdinv))


def test_leaves_of_order_256_declined():
    """Leaves of 256 or more samples (recursive leaf factor in the
    reference) are declined with nothing written; the oracle does not
    restate them either."""
    nl = 2
    prob = workloads.sampled_problem(nl, nblocks=2, hmin=129, hmax=129, t=4, seed=1, lift=False)
    assert prob["blocks"][0]["k"] == 257
    border = prob["border"].copy()
    import pytest
    with pytest.raises(sdpx_gpu.DeviceError, match="declined"):
        sdpx_gpu.sampled_arrow(prob, nl, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    ok, _, _, _ = oracle_sampled_arrow(prob, nl)
    assert ok is None
    assert np.all(mp.same_bits(prob["border"], border))


def _exact_factored_solve(outs, t, border, bdinv, rhs, nrhs, plant=None):
    """The arrow solve's formulas in exact rational arithmetic (no
    rounding) on the computed factors: w = L^-1 b, v = w D^-1, c' = c -
    sum Y^T v, y = L_b^-T D_b^-1 L_b^-1 c', x = L^-T D^-1 (w - Y y)."""
    import pins

    def lower(rec, n):
        v = pins.dec(rec)
        return [[v[i + j * n] for j in range(n)] for i in range(n)]

    def fwd(Lm, b):
        x = list(b)
        for i in range(len(x)):
            x[i] -= sum(Lm[i][k] * x[k] for k in range(i))
        return x

    def bwd(Lm, b):
        x = list(b)
        for i in reversed(range(len(x))):
            x[i] -= sum(Lm[k][i] * x[k] for k in range(i + 1, len(x)))
        return x

    ks = [d.size for _, d, _ in outs]
    n = sum(ks) + t
    rh = pins.dec(rhs)
    Lb, db = lower(border, t), pins.dec(bdinv)
    facs = [(lower(l, K), pins.dec(d), pins.dec(y)) for (l, d, y), K in zip(outs, ks)]
    out = []
    for c in range(nrhs):
        col = rh[c * n:(c + 1) * n]
        ws, vs, off = [], [], 0
        for (Lm, d, _), K in zip(facs, ks):
            w = fwd(Lm, col[off:off + K])
            ws.append(w)
            vs.append([w[i] * (1 if plant == "coupling without D^-1" else d[i]) for i in range(K)])
            off += K
        cc = list(col[off:])
        for (Lm, d, y), K, v in zip(facs, ks, vs):
            for j in range(t):
                cc[j] -= sum(y[i + j * K] * v[i] for i in range(K))
        yv = fwd(Lb, cc)
        yv = bwd(Lb, [yv[j] * db[j] for j in range(t)])
        xs = []
        for (Lm, d, y), K, w in zip(facs, ks, ws):
            u = [(w[i] - sum(y[i + j * K] * yv[j] for j in range(t))) * d[i] for i in range(K)]
            xs += bwd(Lm, u)
        out += xs + yv
    return out


def test_oracle_arrow_solve_is_the_factored_solve():
    """The oracle's arrow solve (FMA sweeps and exact coupling dots, each
    rounded) agrees with the same formulas evaluated without rounding to
    within 2^-150 relative at 256 bits (the leaves' weights down to 1e-6
    cost some 60 bits of conditioning); a wrong index, operand or order of
    the block elimination misses by far more, as a planted one shows."""
    from fractions import Fraction

    from helpers import oracle_sampled_arrow_solve

    nl, nrhs = 4, 2
    prob = workloads.sampled_problem(nl, nblocks=3, hmin=4, hmax=7, t=9, seed=13)
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl)
    assert ok
    n = sum(b["k"] for b in prob["blocks"]) + prob["t"]
    rng = np.random.default_rng(2)
    rhs = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    x = oracle_sampled_arrow_solve(outs, prob["t"], border, bdinv, rhs, nrhs)
    import pins

    got = pins.dec(x)
    ref = _exact_factored_solve(outs, prob["t"], border, bdinv, rhs, nrhs)
    scale = max(abs(v) for v in ref)
    assert scale > 0
    worst = max(abs(a - b) for a, b in zip(got, ref)) / scale
    assert worst < Fraction(1, 2 ** 150), float(worst)
    bad = _exact_factored_solve(outs, prob["t"], border, bdinv, rhs, nrhs, plant="coupling without D^-1")
    assert max(abs(a - b) for a, b in zip(got, bad)) / scale > Fraction(1, 2 ** 20)


def test_cpu_residue_products_keep_the_arrow_bits():
    """The sampled arrow restatement with its products on the CPU residue
    product (as the CPU baseline runs it) gives the exact-dot restatement's
    bits: leaf factors, Y, the border factor."""
    from helpers import oracle_set_rns_products

    nl = 4
    prob = workloads.sampled_problem(nl, nblocks=4, hmin=12, hmax=30, t=300, seed=21)
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl)
    oracle_set_rns_products(True)
    try:
        ok2, border2, bdinv2, outs2 = oracle_sampled_arrow(prob, nl)
    finally:
        oracle_set_rns_products(False)
    assert ok and ok2
    assert np.all(mp.same_bits(border2, border)) and np.all(mp.same_bits(bdinv2, bdinv))
    for a, b in zip(outs, outs2):
        for x, y in zip(a, b):
            assert np.all(mp.same_bits(x, y))
