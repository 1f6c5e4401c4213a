"""Float64 Schur assembly, dense Cholesky and solves on the GPU.

Float64 parity is by tolerance (the reference's BLAS/LAPACK path fixes no
summation order, so bits are unpinned): the factor reproduces A within
1e-12 relative, solves agree with numpy/LAPACK within 1e-12 relative on
well-conditioned systems, and the Schur matrix agrees with a float128
(longdouble) evaluation of the reference's formula within 1e-12 relative.
Known answers: the reference's Cholesky test data (cholesky.rs:307-430,
including logdet 5.69035945432406) and the dynamic pivot rule
(dense_block.rs:1108-1124).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd(n, rng, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    return (q * s) @ q.T


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 1000, 1037])
def test_cholesky_and_solve(gpu, n):
    rng = np.random.default_rng(n)
    a = spd(n, rng)
    f = gpu.DenseFactor(n)
    assert f.refactor(np.triu(a) + np.tril(rng.standard_normal((n, n)), -1))  # lower ignored
    for nrhs in (1, 16, 64):
        b = rng.standard_normal((n, nrhs))
        x = f.solve_many(b)
        ref = np.linalg.solve(a, b)
        assert np.linalg.norm(x - ref) <= 1e-12 * np.linalg.cond(a) * np.linalg.norm(ref) + 1e-300
        res = np.linalg.norm(a @ x - b) / (np.linalg.norm(a) * np.linalg.norm(x))
        assert res < 1e-14
    f.close()


def test_reference_known_answers(gpu):
    cases = [
        ([[4, 1], [1, 3]], [[2, 3], [1, 2]], [[9, 14], [5, 9]]),
        ([[8, -2, 4], [-2, 12, 2], [4, 2, 6]], [[1, 2], [3, 4], [5, 6]],
         [[22, 32], [44, 56], [40, 52]]),
        ([[10, 2, 3, 1], [2, 8, 0, 3], [3, 0, 6, 2], [1, 3, 2, 9]],
         [[1, 2], [2, 3], [3, 1], [4, 2]], [[27, 31], [30, 34], [29, 16], [49, 31]]),
    ]
    for s, x, b in cases:
        s, x, b = (np.array(v, dtype=float) for v in (s, x, b))
        f = gpu.DenseFactor(s.shape[0])
        assert f.refactor(s)
        assert np.abs(f.solve_many(b) - x).max() <= 1e-12
        L = f.factor()
        assert np.abs(L @ L.T - s).max() < 1e-12 * np.abs(s).max()
        if s.shape[0] == 3:
            # CholeskyEngine::logdet known answer (cholesky.rs:416-428)
            assert abs(2 * np.log(np.diag(L)).sum() - 5.69035945432406) < 1e-10
        f.close()


def test_indefinite_reports_pivot_column(gpu):
    for n, bad in ((4, 3), (130, 77), (300, 250)):
        rng = np.random.default_rng(bad)
        a = spd(n, rng, cond=10)
        d = np.ones(n)
        d[bad - 1] = -1.0
        q = np.eye(n)
        a = np.diag(d) if n == 4 else a
        if n != 4:
            # make the leading minor of order `bad` indefinite
            a = a.copy()
            a[bad - 1, bad - 1] = -1e3
        f = gpu.DenseFactor(n)
        assert not f.refactor(a)
        if n == 4:
            assert f.info == bad
        else:
            assert 1 <= f.info <= bad
        f.close()
        del q


def test_dynamic_pivot_rule(gpu):
    # dense_block_replaces_tiny_pivot: a 1e-30 pivot becomes delta
    f = gpu.DenseFactor(1)
    assert not f.refactor(np.array([[-1e-30]]))  # without the rule: rejected
    f.set_dynamic_pivot(1e-13, 2e-7)
    assert f.refactor(np.array([[1e-30]]))
    x = f.solve_many(np.array([[1.0]]))
    assert x[0, 0] == pytest.approx(1.0 / 2e-7, rel=1e-15)
    # a pivot that cancels to rounding level relative to its diagonal
    # (|d| <= 64 eps n |a_kk|) is replaced; a negative pivot of the size of
    # its own diagonal is indefinite and rejected (dense_block.rs:341-345)
    f2 = gpu.DenseFactor(2)
    f2.set_dynamic_pivot(1e-13, 2e-7)
    assert f2.refactor(np.array([[3.0, 1.0], [1.0, 1.0 / 3.0]]))
    f2.close()
    assert not f.refactor(np.array([[-1e-30]]))
    assert not f.refactor(np.array([[-1.0]]))
    f.close()


def svec(mat):
    n = mat.shape[0]
    iu = np.triu_indices(n)
    order = np.lexsort((iu[0], iu[1]))  # column-major upper order
    i, j = iu[0][order], iu[1][order]
    v = mat[i, j].copy()
    v[i != j] *= np.sqrt(2.0)
    return v


@pytest.mark.parametrize("n,m", [(50, 100), (17, 40), (200, 1000)])
def test_schur_psd_matches_reference_formula(gpu, n, m):
    rng = np.random.default_rng(n + m)
    g = rng.standard_normal((n, n))
    rinv = np.linalg.cholesky(np.eye(n) + g @ g.T / n).T  # any invertible factor
    mats = []
    for _ in range(m):
        x = rng.standard_normal((n, n))
        mats.append((x + x.T) / 2)
    a_svec = np.stack([svec(x) for x in mats], axis=1)
    S = gpu.schur_psd_f64(rinv, a_svec)
    # reference formula (psd.rs): S_ab = svec(A_a) . svec(Ginv A_b Ginv), Ginv = Rinv^T Rinv,
    # evaluated in extended precision
    ginv = (rinv.T.astype(np.longdouble) @ rinv.astype(np.longdouble))
    V = np.stack([svec(ginv @ x.astype(np.longdouble) @ ginv) for x in mats], axis=1)
    ref = (a_svec.astype(np.longdouble).T @ V)
    iu = np.triu_indices(m)
    err = np.abs(S[iu] - ref[iu].astype(np.float64)).max()
    assert err <= 1e-12 * float(np.abs(ref).max())


def test_ipm_dense_sdp_n50_m100_cpu_and_gpu(gpu):
    """The plain Float64 SDP (one PSD block n = 50, m = 100 equality
    constraints, X = I strictly feasible) solved to 1e-8 with the Schur
    build + factor + solves on the CPU (numpy / LAPACK) and on the GPU: both
    solve, with the same optimal value and multipliers."""
    from sdpx_gpu import ipm

    prob = ipm.dense_problem(50, 100, seed=0)
    Xc, yc, Sc, ic = ipm.solve(prob, backend="cpu", tol=1e-8)
    Xg, yg, Sg, ig = ipm.solve(prob, backend="gpu", tol=1e-8)
    assert ic["status"] == "solved" and ig["status"] == "solved", (ic["status"], ig["status"])
    assert abs(ig["pobj"] - ic["pobj"]) <= 1e-7 * (1 + abs(ic["pobj"]))
    assert np.linalg.norm(yg - yc) <= 1e-5 * (1 + np.linalg.norm(yc))
    # strict complementarity at the solution: <X, S> small against |C|
    assert np.sum(Xg * Sg) <= 1e-6 * (1 + np.linalg.norm(prob["C"]))


def test_condensed_assembly_several_cones(gpu):
    """P + three PSD blocks (uneven orders, overlapping column subsets) in
    cone order + orthant rows' Gram, against the same sums in extended
    precision, within 1e-12 relative (CondensedKKT::assemble,
    kkt/condensed/mod.rs:1962-2147)."""
    from sdpx_gpu import ipm

    rng = np.random.default_rng(12)
    m = 60
    blocks, ref = [], np.zeros((m, m), dtype=np.longdouble)
    # P: a sparse symmetric pattern, stored upper CSC
    Pd = np.zeros((m, m))
    for _ in range(80):
        i, j = sorted(rng.integers(0, m, 2))
        Pd[i, j] = Pd[j, i] = rng.standard_normal()
    colptr, rowval, nzval = [0], [], []
    for j in range(m):
        for i in range(j + 1):
            if Pd[i, j] != 0:
                rowval.append(i)
                nzval.append(Pd[i, j])
        colptr.append(len(rowval))
    ref += np.triu(Pd).astype(np.longdouble)
    for n, k in ((8, 25), (15, 30), (5, 60)):
        cols = np.sort(rng.choice(m, k, replace=False))
        rinv = np.eye(n) + rng.standard_normal((n, n)) / (4 * np.sqrt(n))
        A = [(lambda g: g + g.T)(rng.standard_normal((n, n))) for _ in range(k)]
        a_svec = np.stack([ipm.svec(a) for a in A], axis=1)
        blocks.append((rinv, a_svec, cols))
        W = (rinv.T @ rinv).astype(np.longdouble)
        for a in range(k):
            WAW = W @ A[a].astype(np.longdouble) @ W
            for b2 in range(k):
                i, j = min(cols[a], cols[b2]), max(cols[a], cols[b2])
                if a <= b2:
                    ref[i, j] += np.sum(A[b2].astype(np.longdouble) * WAW)
    r = 20
    B = rng.standard_normal((r, m))
    w = np.exp(rng.standard_normal(r))
    Bl = B.astype(np.longdouble)
    ref += np.triu(Bl.T @ (w[:, None].astype(np.longdouble) * Bl))
    S = gpu.condensed_assemble_f64(m, (colptr, rowval, nzval), blocks, B, w)
    scale = float(np.abs(ref).max())
    err = np.abs(np.triu(S).astype(np.longdouble) - ref).max() / scale
    assert float(err) <= 1e-12, float(err)


def test_condensed_solve_recovers_the_unreduced_system(gpu):
    """CondensedKKT::solve_raw on the GPU (prepare_rhs, the reduced factor
    solve, recover_rhs; kkt/condensed/mod.rs:2416-2610) for P, three PSD
    blocks (one with no constraint columns) and orthant rows: (x, z) solve
    the unreduced system P x + A^T z = bx, A x - H z = bz, with H^-1 the
    cones' inverse scalings (svec(W smat(.) W), W = Rinv^T Rinv; w o .), in
    extended precision to 1e-11 relative; z also against the recovery
    formula and x against numpy's solve of the assembled system."""
    from sdpx_gpu import ipm

    rng = np.random.default_rng(21)
    m = 50
    Pd = np.eye(m) * 0.5
    for _ in range(40):
        i, j = sorted(rng.integers(0, m, 2))
        v = 0.1 * rng.standard_normal()
        Pd[i, j] += v
        if i != j:
            Pd[j, i] += v
    Pd += m * 0.02 * np.eye(m)
    colptr, rowval, nzval = [0], [], []
    for j in range(m):
        for i in range(j + 1):
            if Pd[i, j] != 0:
                rowval.append(i)
                nzval.append(Pd[i, j])
        colptr.append(len(rowval))
    blocks, mats = [], []
    for n, k in ((7, 20), (12, 35), (4, 0)):
        cols = np.sort(rng.choice(m, k, replace=False))
        rinv = np.eye(n) + rng.standard_normal((n, n)) / (4 * np.sqrt(n))
        A = [(lambda g: g + g.T)(rng.standard_normal((n, n))) for _ in range(k)]
        a_svec = np.stack([ipm.svec(a) for a in A], axis=1) if k else np.zeros((n * (n + 1) // 2, 0))
        blocks.append((rinv, a_svec, cols))
        mats.append((n, rinv, a_svec, cols))
    r = 15
    B = rng.standard_normal((r, m))
    w = np.exp(rng.standard_normal(r))
    ssum = sum(n * (n + 1) // 2 for n, *_ in mats)
    bx = rng.standard_normal(m)
    bz = rng.standard_normal(ssum + r)
    x, z = gpu.condensed_solve_f64(m, (colptr, rowval, nzval), blocks, B, w, bx, bz)

    ld = np.longdouble
    hinv = lambda rinv, v: ipm.svec((rinv.T @ rinv) @ ipm.smat(v, rinv.shape[0]) @ (rinv.T @ rinv))
    xl, zl = x.astype(ld), z.astype(ld)
    r1 = Pd.astype(ld) @ xl - bx.astype(ld)
    scale = np.abs(Pd).max() * np.abs(x).max() + np.abs(bx).max()
    off = 0
    for n, rinv, a_svec, cols in mats:
        s = n * (n + 1) // 2
        zb = z[off:off + s]
        if len(cols):
            r1[cols] += a_svec.astype(ld).T @ zl[off:off + s]
        ax = a_svec @ x[cols] if len(cols) else np.zeros(s)
        want = hinv(rinv, ax - bz[off:off + s])
        assert np.abs(zb - want).max() <= 1e-11 * max(1.0, np.abs(want).max())
        scale += np.abs(a_svec).max(initial=0) * np.abs(zb).max() * max(1, len(cols))
        off += s
    zo = z[off:]
    assert np.abs(zo - w * (B @ x - bz[off:])).max() <= 1e-11 * max(1.0, np.abs(zo).max())
    r1 += B.astype(ld).T @ zl[off:]
    assert float(np.abs(r1).max()) <= 1e-11 * scale, float(np.abs(r1).max() / scale)
    S = gpu.condensed_assemble_f64(m, (colptr, rowval, nzval), blocks, B, w)
    S = np.triu(S) + np.triu(S, 1).T
    xr = bx.copy()
    off = 0
    for n, rinv, a_svec, cols in mats:
        s = n * (n + 1) // 2
        if len(cols):
            xr[cols] += a_svec.T @ hinv(rinv, bz[off:off + s])
        off += s
    xr += B.T @ (w * bz[off:])
    xn = np.linalg.solve(S, xr)
    assert np.abs(x - xn).max() <= 1e-9 * np.abs(xn).max()


def test_conic_json_problem_through_the_condensed_solve(gpu):
    """A conic JSON problem (PSD triangle and nonnegative cones, read by
    sdpx_gpu.conic_json) mapped onto the condensed assembly, solved on the
    GPU (assembly, factor, prepare_rhs, recover_rhs): with the problem's own
    rows A_K in cone order, P x + A_K^T z = bx to 1e-11 relative."""
    import json

    from sdpx_gpu import conic_json as cj

    rng = np.random.default_rng(8)
    cones = [("psd", 4), ("nonneg", 6), ("psd", 3)]
    m = sum(cj.cone_size(c) for c in cones)
    n = 30
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.5)
    Pd = np.eye(n) + 0.05 * rng.standard_normal((n, n))
    Pd = Pd @ Pd.T

    def csc(M):
        cp, rv, nz = [0], [], []
        for j in range(M.shape[1]):
            for i in range(M.shape[0]):
                if M[i, j] != 0:
                    rv.append(i)
                    nz.append(float(M[i, j]))
            cp.append(len(rv))
        return {"m": M.shape[0], "n": M.shape[1], "colptr": cp, "rowval": rv, "nzval": nz}

    doc = {"P": csc(np.triu(Pd)), "q": [0.0] * n, "A": csc(A), "b": [0.0] * m,
           "cones": [cj._cone_json(c) for c in cones]}
    pr = cj.read_conic_json(json.dumps(doc))
    rinvs = [np.eye(4) + 0.2 * rng.standard_normal((4, 4)), np.eye(3) + 0.2 * rng.standard_normal((3, 3))]
    w = np.exp(rng.standard_normal(6))
    p_csc, blocks, B, ww, retained = cj.condensed_inputs(pr, rinvs, w)
    assert retained == []
    bx = rng.standard_normal(n)
    bz_cone = rng.standard_normal(m)  # eliminated rows in cone order
    psd_rows = [slice(0, 10), slice(16, 22)]
    bz = np.concatenate([bz_cone[psd_rows[0]], bz_cone[psd_rows[1]], bz_cone[10:16]])
    x, z = gpu.condensed_solve_f64(n, p_csc, blocks, B, ww, bx, bz)
    z_cone = np.empty(m)
    z_cone[psd_rows[0]], z_cone[psd_rows[1]], z_cone[10:16] = z[:10], z[10:16], z[16:]
    # the library's svec rows are the cone's rows over the touched columns:
    # A_K x restricted to those columns is A_K x itself
    ld = np.longdouble
    r1 = Pd.astype(ld) @ x.astype(ld) + A.astype(ld).T @ z_cone.astype(ld) - bx.astype(ld)
    scale = np.abs(Pd).max() * np.abs(x).max() + np.abs(A).max() * np.abs(z).max() * m + np.abs(bx).max()
    assert float(np.abs(r1).max()) <= 1e-11 * scale
