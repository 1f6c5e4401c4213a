"""Adjoint (differentiation) of the Float64 PSD Schur path: gradients of
y* and of the optimal objective with respect to b and C through one
transposed-KKT solve with the GPU factor, checked against the CPU adjoint
solve (numpy/LAPACK: the GPU solution's normwise backward error is at the
binary64 level, and the two solutions agree within 1e-10 relative plus
16 eps cond(M), the forward error two backward-stable solves may differ by
at the ill-conditioned central-path point), against central finite differences
of full IPM re-solves (step 1e-6), and through a torch.autograd.Function
binding (test-only).  The reference has no differentiation: FD and the CPU
adjoint are the pins."""
import numpy as np
import pytest

import sdpx_gpu
from sdpx_gpu import ipm

pytestmark = pytest.mark.gpu


EPS = np.finfo(np.float64).eps


def check_solve(M, v_gpu, v_cpu, rhs):
    """Backward error of the GPU solve, then its forward distance to the
    CPU solve within what the conditioning of M allows."""
    res = np.abs(M @ v_gpu - rhs).max(axis=0)
    back = res / (np.abs(M).sum(axis=1).max() * np.abs(v_gpu).max(axis=0) + np.abs(rhs).max(axis=0))
    assert back.max() <= 64 * M.shape[0] * EPS, back.max()
    cond = np.linalg.cond(M)
    tol = (1e-10 + 16 * EPS * cond) * np.abs(v_cpu).max()
    print(f"cond(M) {cond:.3e}, backward {back.max():.2e}, forward {np.abs(v_gpu - v_cpu).max():.2e}")
    assert np.abs(v_gpu - v_cpu).max() <= tol


def check_congruence(prob, W, v, cg, cols):
    """C-gradient W A*(v) W of the GPU's own v, within the forward error
    bound of the matrix products (64 n eps |W| |A*(v)| |W|)."""
    n = prob["n"]
    for k in cols:
        a = ipm.smat(prob["a_svec"] @ v[:, k], n)
        ref = ipm.svec(W @ a @ W)
        bound = 64 * n * EPS * np.abs(W) @ np.abs(a) @ np.abs(W)
        assert np.all(np.abs(cg[:, k] - ref) <= ipm.svec(bound) + 1e-300)


def cpu_adjoint(prob, W, ybar):
    m, n = prob["m"], prob["n"]
    A = prob["A"]
    M = np.array([[np.sum(A[i] * (W @ A[j] @ W)) for j in range(m)] for i in range(m)])
    v = np.linalg.solve(M, ybar)
    cg = np.stack([ipm.svec(W @ ipm.smat(prob["a_svec"] @ v[:, k], n) @ W)
                   for k in range(v.shape[1])], axis=1)
    return v, cg


MU = 1e-6


def solve(prob, mu=MU):
    """The central-path point X S = mu I (the adjoint is its exact
    derivative, so central differences pin it at step 1e-6)."""
    X, y, S, info = ipm.solve(prob, backend="gpu", target_mu=mu, max_iter=200)
    assert info["status"] == "centred", info["status"]
    return X, y, S, info


def test_adjoint_matches_cpu_and_fd(gpu):
    prob = ipm.dense_problem(20, 30, seed=3)
    X, y, S, info = solve(prob)
    rng = np.random.default_rng(0)
    ybar = rng.standard_normal((prob["m"], 8))
    # the adjoint linearizes at the final iterate: its scaling and Schur factor
    W, rinv = info["W"], info["rinv"]
    M = sdpx_gpu.schur_psd_f64(rinv, prob["a_svec"])
    f = sdpx_gpu.DenseFactor(prob["m"])
    assert f.refactor(np.triu(M) + np.triu(M, 1).T)
    bg, cg = sdpx_gpu.adjoint_psd_f64(f, rinv, prob["a_svec"], ybar)
    vb, vc = cpu_adjoint(prob, W, ybar)
    Ms = np.triu(M) + np.triu(M, 1).T
    check_solve(Ms, bg, vb, ybar)
    check_congruence(prob, W, bg, cg, range(bg.shape[1]))
    # central finite differences of y* along random directions in b and C
    eps = 1e-6
    for k in range(2):
        db = rng.standard_normal(prob["m"])
        pp, pm = dict(prob), dict(prob)
        pp["b"], pm["b"] = prob["b"] + eps * db, prob["b"] - eps * db
        yp = solve(pp)[1]
        ym = solve(pm)[1]
        fd = ybar[:, k] @ (yp - ym) / (2 * eps)
        assert abs(fd - bg[:, k] @ db) <= 1e-4 * (1 + abs(fd))
        dc = rng.standard_normal((prob["n"], prob["n"]))
        dc = (dc + dc.T) / 2
        pp, pm = dict(prob), dict(prob)
        pp["C"], pm["C"] = prob["C"] + eps * dc, prob["C"] - eps * dc
        yp = solve(pp)[1]
        ym = solve(pm)[1]
        fd = ybar[:, k] @ (yp - ym) / (2 * eps)
        assert abs(fd - cg[:, k] @ ipm.svec(dc)) <= 1e-4 * (1 + abs(fd))
    # objective at the solution: d p*/d b = y*, d p*/d C = X* (envelope
    # theorem), checked by central differences of solves to 1e-8
    sol = ipm.solve(prob, backend="gpu", tol=1e-8)
    db = rng.standard_normal(prob["m"])
    pp, pm = dict(prob), dict(prob)
    pp["b"], pm["b"] = prob["b"] + 1e-4 * db, prob["b"] - 1e-4 * db
    fd = (ipm.solve(pp, backend="gpu", tol=1e-8)[3]["dobj"]
          - ipm.solve(pm, backend="gpu", tol=1e-8)[3]["dobj"]) / 2e-4
    assert abs(fd - sol[1] @ db) <= 1e-4 * (1 + abs(fd))


def test_adjoint_batch_64_at_n200_m1000(gpu):
    prob = ipm.dense_problem(200, 1000, seed=4)
    X, y, S, info = solve(prob)
    W, rinv = info["W"], info["rinv"]
    M = sdpx_gpu.schur_psd_f64(rinv, prob["a_svec"])
    M = np.triu(M) + np.triu(M, 1).T
    f = sdpx_gpu.DenseFactor(prob["m"])
    assert f.refactor(M)
    ybar = np.random.default_rng(1).standard_normal((prob["m"], 64))
    bg, cg = sdpx_gpu.adjoint_psd_f64(f, rinv, prob["a_svec"], ybar)
    v = np.linalg.solve(M, ybar)
    check_solve(M, bg, v, ybar)
    check_congruence(prob, W, bg, cg, (0, 5, 63))


def test_torch_autograd_binding(gpu):
    import torch
    from adjoint_torch import SdpSolve

    prob = ipm.dense_problem(12, 15, seed=9)
    b = torch.tensor(prob["b"], dtype=torch.float64, requires_grad=True)
    c = torch.tensor(ipm.svec(prob["C"]), dtype=torch.float64, requires_grad=True)
    y = SdpSolve.apply(b, c, prob)
    w = torch.tensor(np.random.default_rng(2).standard_normal(prob["m"]))
    (y @ w).backward()
    eps = 1e-6
    db = np.random.default_rng(3).standard_normal(prob["m"])
    pp, pm = dict(prob), dict(prob)
    pp["b"], pm["b"] = prob["b"] + eps * db, prob["b"] - eps * db
    fd = w.numpy() @ (solve(pp)[1] - solve(pm)[1]) / (2 * eps)
    assert abs(fd - b.grad.numpy() @ db) <= 1e-4 * (1 + abs(fd))
    assert c.grad is not None and c.grad.shape == c.shape
