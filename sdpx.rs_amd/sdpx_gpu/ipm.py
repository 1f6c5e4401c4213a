"""A compact primal-dual interior-point driver for dense SDPs whose per
iteration Schur build + factor + solves run on the GPU hot path.

This stands in for the reference's HSD loop (crates/solver/src/solver/core/
solver.rs:285-506), which is the caller of the path and not itself rebuilt
here.  It poses the problem as the reference does for the Float64 configs:
free variables y and one PSD slack S = C - sum_i y_i A_i, i.e.

    primal  min <C, X>   s.t. <A_i, X> = b_i, X PSD
    dual    max b^T y    s.t. C - sum_i y_i A_i = S PSD

and takes Nesterov-Todd directions with a Mehrotra-style centring.  The
condensed Schur matrix is M_ij = <A_i, W A_j W> with the NT scaling W; with
W = Rinv^T Rinv this is exactly the reference's S_ab = <A_a, Ginv A_b Ginv>
(kkt/condensed/psd.rs:1299-1456, kkt.rs:1119-1127).

backend "gpu": the NT scaling (sdpx_gpu_psd_nt_scaling_f64_dev),
sdpx_gpu_schur_psd_f64 and the dense GPU factor / solve;
backend "cpu": the same formulas in numpy / LAPACK.
"""
import time

import numpy as np

from . import DenseFactor, psd_nt_scaling_f64, schur_psd_f64


def svec(mat):
    n = mat.shape[0]
    iu = np.triu_indices(n)
    order = np.lexsort((iu[0], iu[1]))  # column-major upper order
    i, j = iu[0][order], iu[1][order]
    v = mat[i, j].copy()
    v[i != j] *= np.sqrt(2.0)
    return v


def smat(v, n):
    iu = np.triu_indices(n)
    order = np.lexsort((iu[0], iu[1]))
    i, j = iu[0][order], iu[1][order]
    out = np.zeros((n, n))
    vals = v.copy()
    vals[i != j] /= np.sqrt(2.0)
    out[i, j] = vals
    out[j, i] = vals
    return out


def dense_problem(n, m, seed):
    """A_i symmetric with i.i.d. N(0,1) entries, b_i = <A_i, I> (X = I is
    strictly feasible), C = sum y0_i A_i + I + G G^T / n."""
    rng = np.random.default_rng(seed)
    A = []
    for _ in range(m):
        x = rng.standard_normal((n, n))
        A.append(np.triu(x) + np.triu(x, 1).T)
    b = np.array([np.trace(a) for a in A])
    y0 = rng.standard_normal(m)
    g = rng.standard_normal((n, n))
    C = sum(yi * a for yi, a in zip(y0, A)) + np.eye(n) + g @ g.T / n
    a_svec = np.stack([svec(a) for a in A], axis=1)
    return {"n": n, "m": m, "A": A, "a_svec": np.asfortranarray(a_svec), "b": b, "C": C}


def _nt_scaling(X, S):
    """W with W S W = X and a factor Rinv with W = Rinv^T Rinv."""
    lx = np.linalg.cholesky(X)
    ls = np.linalg.cholesky(S)
    u, sv, vt = np.linalg.svd(ls.T @ lx)
    # W = L_X V diag(sv)^-1 V^T L_X^T  (Todd-Toh-Tutuncu)
    r = lx @ vt.T @ np.diag(sv ** -0.5)  # W = R R^T
    W = r @ r.T
    return W, r.T  # Rinv := R^T so that W = Rinv^T Rinv


class Stats(dict):
    pass


def solve(prob, backend="gpu", tol=1e-8, max_iter=80, verbose=False, target_mu=None, centre_tol=1e-9):
    """Returns (X, y, S, info).  With target_mu, returns instead the point of
    the central path X S = target_mu I (Newton centring steps once mu is
    near it, until |X S - mu I| / (mu sqrt n) < centre_tol); the adjoint of
    the Schur solve is the exact derivative of that point, so finite
    differences of it pin the adjoint at any step size."""
    n, m = prob["n"], prob["m"]
    A, a_svec, b, C = prob["A"], prob["a_svec"], prob["b"], prob["C"]

    def Aop(X):
        return a_svec.T @ svec(X)

    def Aadj(y):
        return smat(a_svec @ y, n)

    X = np.eye(n)
    y = np.zeros(m)
    S = np.eye(n)
    factor = DenseFactor(m) if backend == "gpu" else None
    info = Stats(iterations=0, schur_s=0.0, factor_s=0.0, solve_s=0.0, status="max_iter")
    for it in range(max_iter):
        rp = b - Aop(X)
        Rd = C - Aadj(y) - S
        mu = np.trace(X @ S) / n
        pobj, dobj = np.sum(C * X), b @ y
        gap = abs(pobj - dobj) / (1 + abs(pobj) + abs(dobj))
        pinf = np.linalg.norm(rp) / (1 + np.linalg.norm(b))
        dinf = np.linalg.norm(Rd) / (1 + np.linalg.norm(C))
        if verbose:
            print(f"{it:3d} {pobj:+.10e} {dobj:+.10e} gap {gap:.2e} p {pinf:.2e} d {dinf:.2e}")
        centring = target_mu is not None and mu < 10 * target_mu
        if target_mu is not None:
            cres = np.linalg.norm(X @ S - target_mu * np.eye(n)) / (target_mu * np.sqrt(n))
            if centring and cres < centre_tol and pinf < 1e-10 and dinf < 1e-12:
                info["status"] = "centred"
                break
        elif gap < tol and pinf < tol and dinf < tol:
            info["status"] = "solved"
            break
        if backend == "gpu":
            r, _, _, W, status = psd_nt_scaling_f64([(X, S)])[0]
            if status != 0:
                info["status"] = "scaling_failed"
                break
            rinv = r.T
        else:
            W, rinv = _nt_scaling(X, S)
        t0 = time.perf_counter()
        if backend == "gpu":
            M = schur_psd_f64(rinv, a_svec)
            M = np.triu(M) + np.triu(M, 1).T
        else:
            Ms = np.stack([svec(rinv @ A[i] @ rinv.T) for i in range(m)], axis=1)
            M = Ms.T @ Ms
        t1 = time.perf_counter()
        if backend == "gpu":
            ok = factor.refactor(M)
        else:
            try:
                Lm = np.linalg.cholesky(M)
                ok = True
            except np.linalg.LinAlgError:
                ok = False
        t2 = time.perf_counter()
        if not ok:
            info["status"] = "factor_failed"
            break

        def msolve(r):
            if backend == "gpu":
                return factor.solve_many(r.reshape(m, -1)).reshape(r.shape)
            return np.linalg.solve(Lm.T, np.linalg.solve(Lm, r))

        Sinv = np.linalg.inv(S)
        WRdW = W @ Rd @ W

        def direction(Rc):
            rhs = rp - Aop(Rc - WRdW)
            dy = msolve(rhs)
            dS = Rd - Aadj(dy)
            dX = Rc - W @ dS @ W
            dX = (dX + dX.T) / 2
            return dX, dy, dS

        def step(Z, dZ):
            lz = np.linalg.cholesky(Z)
            li = np.linalg.inv(lz)
            ev = np.linalg.eigvalsh(li @ dZ @ li.T)
            lo = ev.min()
            return 1.0 if lo >= 0 else min(1.0, -1.0 / lo)

        if centring:
            dX, dy, dS = direction(target_mu * Sinv - X)
            ap, ad = min(1.0, 0.95 * step(X, dX) if step(X, dX) < 1 else 1.0), 1.0
            ad = min(1.0, 0.95 * step(S, dS) if step(S, dS) < 1 else 1.0)
        else:
            dXa, dya, dSa = direction(-X)
            ap, ad = step(X, dXa), step(S, dSa)
            mua = np.trace((X + ap * dXa) @ (S + ad * dSa)) / n
            sigma = min(1.0, (mua / mu) ** 3)
            if target_mu is not None:
                sigma = max(sigma, target_mu / mu)
            dX, dy, dS = direction(sigma * mu * Sinv - X)
            ap, ad = 0.95 * step(X, dX), 0.95 * step(S, dS)
            ap, ad = min(ap, 1.0), min(ad, 1.0)
        # keep both iterates numerically positive definite (backtracking)
        for _ in range(30):
            try:
                np.linalg.cholesky(X + ap * dX)
                break
            except np.linalg.LinAlgError:
                ap *= 0.8
        for _ in range(30):
            try:
                np.linalg.cholesky(S + ad * dS)
                break
            except np.linalg.LinAlgError:
                ad *= 0.8
        X = X + ap * dX
        X = (X + X.T) / 2
        y = y + ad * dy
        S = S + ad * dS
        S = (S + S.T) / 2
        t3 = time.perf_counter()
        info["iterations"] = it + 1
        info["schur_s"] += t1 - t0
        info["factor_s"] += t2 - t1
        info["solve_s"] += t3 - t2
    info["pobj"] = float(np.sum(C * X))
    info["dobj"] = float(b @ y)
    info["factor"] = factor
    info["W"], info["rinv"] = _nt_scaling(X, S)
    return X, y, S, info
