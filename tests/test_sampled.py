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
