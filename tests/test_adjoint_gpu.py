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


def solve(prob, mu=MU, centre_tol=1e-9):
    """The central-path point X S = mu I (the adjoint is its exact
    derivative, so central differences pin it at step 1e-6)."""
    X, y, S, info = ipm.solve(prob, backend="gpu", target_mu=mu, max_iter=300, centre_tol=centre_tol)
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


def cpu_full_adjoint(prob, W, y, X, ybar, xbar, obj):
    """numpy restatement of the full adjoint at the central-path point."""
    n = prob["n"]
    A = prob["A"]
    m = len(A)
    Msch = np.array([[np.sum(A[i] * (W @ A[j] @ W)) for j in range(m)] for i in range(m)])
    r = ybar.shape[1]
    bg = np.zeros((m, r))
    cg = np.zeros((prob["a_svec"].shape[0], r))
    ag = np.zeros((cg.shape[0], m, r))
    for k in range(r):
        Zm = W @ (ipm.smat(xbar[:, k], n) + obj[k] * prob["C"]) @ W
        g = ybar[:, k] + prob["a_svec"].T @ ipm.svec(Zm)
        v = np.linalg.solve(Msch, g)
        U = W @ ipm.smat(prob["a_svec"] @ v, n) @ W
        bg[:, k] = v
        cg[:, k] = ipm.svec(U - Zm + obj[k] * X)
        ag[:, :, k] = np.outer(ipm.svec(Zm - U), y) - np.outer(ipm.svec(X), v)
    return bg, cg, ag


def _functional(prob, ybar, xbar, obj, centre_tol, mu=MU):
    X, y, S, info = solve(prob, mu, centre_tol=centre_tol)
    return ybar @ y + ipm.svec(X) @ xbar + obj * np.sum(prob["C"] * X)


def _perturbed(prob, field, d, eps):
    p = dict(prob)
    if field == "b":
        p["b"] = prob["b"] + eps * d
    elif field == "C":
        p["C"] = prob["C"] + eps * d
    else:  # every A_i along d[i] (symmetric)
        p["A"] = [a + eps * da for a, da in zip(prob["A"], d)]
        p["a_svec"] = np.asfortranarray(np.stack([ipm.svec(a) for a in p["A"]], axis=1))
    return p


@pytest.mark.parametrize("n,m", [(20, 30), (200, 1000)])
def test_full_adjoint_x_y_objective_wrt_b_c_a(gpu, n, m):
    """Gradients of L = ybar . y* + <Xbar, X*> + obj <C, X*> with respect
    to b, C and every A_i from one solve with the GPU factor: against the
    numpy restatement, and against central finite differences (step 1e-6)
    of re-solved central-path points along random directions in b, C and
    A (configs[4] size included).  X* enters L, so the points are centred
    as far as binary64 holds them (|X S - mu I| / (mu sqrt n) below 3e-10
    at n = 20, 1e-9 at n = 200); at n = 200 that leaves a 1e-6 quotient
    some 1e-3 of noise, so the step there is 1e-5."""
    prob = ipm.dense_problem(n, m, seed=3 if n == 20 else 4)
    print("full adjoint", n, m, flush=True)
    ctol = 3e-10 if n == 20 else 1e-9
    X, y, S, info = solve(prob, centre_tol=ctol)
    W, rinv = info["W"], info["rinv"]
    M = sdpx_gpu.schur_psd_f64(rinv, prob["a_svec"])
    f = sdpx_gpu.DenseFactor(m)
    assert f.refactor(np.triu(M) + np.triu(M, 1).T)
    rng = np.random.default_rng(n)
    s = n * (n + 1) // 2
    ybar = rng.standard_normal((m, 2))
    xbar = rng.standard_normal((s, 2))
    obj = np.array([0.5, -1.0])
    want_a = n == 20
    bg, cg, ag = sdpx_gpu.adjoint_psd_full_f64(f, rinv, prob["a_svec"], y, ipm.svec(X), ipm.svec(prob["C"]),
                                               ybar, xbar, obj, want_a=want_a)
    if n == 20:
        cb, cc, ca = cpu_full_adjoint(prob, W, y, X, ybar, xbar, obj)
        scale = lambda a: np.abs(a).max()
        assert np.abs(bg - cb).max() <= 1e-7 * scale(cb)
        assert np.abs(cg - cc).max() <= 1e-7 * scale(cc)
        assert np.abs(ag - ca).max() <= 1e-7 * scale(ca)
    eps = 1e-6 if n == 20 else 1e-5
    for field in ("b", "C", "A"):
        if field == "b":
            d = rng.standard_normal(m)
        elif field == "C":
            d = rng.standard_normal((n, n))
            d = (d + d.T) / 2
        else:
            # every A_i moves; scaled so the whole move is as large as C's
            d = [(lambda g: (g + g.T) / (2 * np.sqrt(m)))(rng.standard_normal((n, n))) for _ in range(m)]
        k = 0
        fp = _functional(_perturbed(prob, field, d, eps), ybar[:, k], xbar[:, k], obj[k], ctol)
        fm = _functional(_perturbed(prob, field, d, -eps), ybar[:, k], xbar[:, k], obj[k], ctol)
        fd = (fp - fm) / (2 * eps)
        if field == "b":
            an = bg[:, k] @ d
        elif field == "C":
            an = cg[:, k] @ ipm.svec(d)
        else:
            dsv = np.stack([ipm.svec(di) for di in d], axis=1)  # s x m
            an = (np.sum(ag[:, :, k] * dsv) if ag is not None else
                  np.sum((np.outer(-cg[:, k] + obj[k] * ipm.svec(X), y) - np.outer(ipm.svec(X), bg[:, k])) * dsv))
        print(field, fd, an, flush=True)
        assert abs(fd - an) <= 1e-4 * (1 + abs(fd)), (field, fd, an)


def test_torch_autograd_full_binding(gpu):
    """torch.autograd through (y*, X*, objective) of (b, C, A): the
    gradient of a random scalarization equals central differences of
    re-solves along random b, C and A directions."""
    import torch
    from adjoint_torch import SdpSolveFull

    prob = ipm.dense_problem(12, 15, seed=9)
    b = torch.tensor(prob["b"], dtype=torch.float64, requires_grad=True)
    c = torch.tensor(ipm.svec(prob["C"]), dtype=torch.float64, requires_grad=True)
    a = torch.tensor(np.array(prob["a_svec"]), dtype=torch.float64, requires_grad=True)
    y, x, obj = SdpSolveFull.apply(b, c, a, prob)
    rng = np.random.default_rng(5)
    wy, wx = torch.tensor(rng.standard_normal(y.shape[0])), torch.tensor(rng.standard_normal(x.shape[0]))
    (y @ wy + x @ wx + 0.3 * obj).backward()
    eps = 1e-6

    def value(bv, cv, av):
        yy, xx, oo = SdpSolveFull.apply(torch.tensor(bv), torch.tensor(cv), torch.tensor(av), prob)
        return float(yy @ wy + xx @ wx + 0.3 * oo)

    b0, c0, a0 = prob["b"], ipm.svec(prob["C"]), np.array(prob["a_svec"])
    for which in ("b", "c", "a"):
        d = {"b": rng.standard_normal(b0.shape), "c": rng.standard_normal(c0.shape),
             "a": rng.standard_normal(a0.shape) / np.sqrt(a0.shape[1])}[which]
        args = lambda sg: [v + sg * eps * d if k == which else v for k, v in (("b", b0), ("c", c0), ("a", a0))]
        fd = (value(*args(1)) - value(*args(-1))) / (2 * eps)
        g = {"b": b.grad, "c": c.grad, "a": a.grad}[which].numpy()
        an = float(np.sum(g * d))
        assert abs(fd - an) <= 1e-4 * (1 + abs(fd)), (which, fd, an)
