"""Nesterov-Todd scaling of PSD blocks on the GPU (sdpx_gpu_psd_nt_scaling_f64_dev)
against the reference's binary64 steps in numpy/LAPACK
(PSDTriangleCone::update_scaling, cones/psdtrianglecone.rs:311-430:
Cholesky of S and Z, SVD of L2^T L1, R = L1 V diag(lambda)^-1/2,
Rinv = diag(lambda)^1/2 (L1^-T V)^T, G = R R^T), within 1e-12 relative on
well-conditioned blocks, plus the properties that fix the scaling whatever
the algorithm: Rinv R = I, R^T Z R = diag(lambda) = Rinv S Rinv^T, and
G Z G = S (G is the NT scaling point W of W Z W = S).  R and Rinv, fixed
only up to the signs of the singular vectors and more sensitive to the gaps
between singular values, are compared at 1e-11."""
import numpy as np
import pytest
import scipy.linalg as sla

import sdpx_gpu

pytestmark = pytest.mark.gpu


def reference_scaling(S, Z):
    """The reference's binary64 formulas, LAPACK doing the factorizations."""
    l1 = np.linalg.cholesky(S)
    l2 = np.linalg.cholesky(Z)
    _, lam, vt = np.linalg.svd(l2.T @ l1)
    v = vt.T
    R = (l1 @ v) * (1.0 / np.sqrt(lam))
    X = sla.solve_triangular(l1, v, lower=True, trans="T")
    Rinv = np.sqrt(lam)[:, None] * X.T
    return R, Rinv, lam, R @ R.T


def spd(n, cond, rng):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.logspace(0, -np.log10(cond), n)) @ q.T


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def check_block(S, Z, got, tol):
    R, Rinv, lam, G, st = got
    assert st == 0
    Rr, Rinvr, lamr, Gr = reference_scaling(S, Z)
    assert np.max(np.abs(lam - lamr)) <= tol * lamr[0]
    assert rel(G, Gr) <= tol
    # singular vectors are fixed up to sign, and more sensitive than the
    # values (by the inverse gap of neighbouring singular values): align
    # the columns of R and compare R and Rinv ten times looser
    sg = np.sign(np.sum(R * Rr, axis=0))
    assert rel(R * sg, Rr) <= 10 * tol
    assert rel(Rinv * sg[:, None], Rinvr) <= 10 * tol
    n = S.shape[0]
    assert np.linalg.norm(Rinv @ R - np.eye(n)) <= tol * n
    assert rel(R.T @ Z @ R, np.diag(lam)) <= tol
    assert rel(Rinv @ S @ Rinv.T, np.diag(lam)) <= tol
    assert rel(G @ Z @ G, S) <= tol


def test_batched_blocks_of_the_sampled_shapes(gpu):
    """40 blocks of orders 10..60 (the sampled workload's PSD blocks) in one
    launch, moderately conditioned iterates."""
    rng = np.random.default_rng(1)
    dims = rng.integers(10, 61, 40)
    pairs = [(spd(n, 1e2, rng), spd(n, 1e2, rng)) for n in dims]
    out = sdpx_gpu.psd_nt_scaling_f64(pairs)
    for (S, Z), got in zip(pairs, out):
        check_block(S, Z, got, 1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 50, 65, 200, 256])
def test_single_block_orders(gpu, n):
    """Orders at the LDS/global boundary (64), the dense configurations'
    n = 50 and n = 200, and the largest admitted order."""
    rng = np.random.default_rng(n)
    S, Z = spd(n, 1e2, rng), spd(n, 1e2, rng)
    check_block(S, Z, sdpx_gpu.psd_nt_scaling_f64([(S, Z)])[0], 1e-12)


def test_ill_conditioned_iterate(gpu):
    """An interior-point iterate near the boundary (condition 1e8 in S and
    Z): the properties hold to working accuracy relative to the conditioning."""
    rng = np.random.default_rng(7)
    S, Z = spd(60, 1e8, rng), spd(60, 1e8, rng)
    R, Rinv, lam, G, st = sdpx_gpu.psd_nt_scaling_f64([(S, Z)])[0]
    assert st == 0
    _, _, lamr, Gr = reference_scaling(S, Z)
    assert np.max(np.abs(lam - lamr) / lamr) <= 1e-7
    assert rel(G, Gr) <= 1e-7
    assert np.linalg.norm(Rinv @ R - np.eye(60)) <= 1e-7


def test_failures_are_per_block(gpu):
    """A block whose S (or Z) is not positive definite reports status 1 (2),
    as the reference's update_scaling returns false; its neighbours in the
    batch are unaffected."""
    rng = np.random.default_rng(3)
    good = (spd(12, 10.0, rng), spd(12, 10.0, rng))
    bad_s = (-spd(9, 10.0, rng), spd(9, 10.0, rng))
    bad_z = (spd(70, 10.0, rng), spd(70, 10.0, rng) - 2.0 * np.eye(70))
    out = sdpx_gpu.psd_nt_scaling_f64([good, bad_s, bad_z, good])
    assert [o[4] for o in out] == [0, 1, 2, 0]
    check_block(*good, out[0], 1e-12)
    check_block(*good, out[3], 1e-12)
