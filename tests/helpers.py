"""Test helpers: the C oracle through ctypes, and conversions between
numpy records and the Python big-integer oracle."""
import ctypes
import os
import subprocess

import numpy as np

import mpexact
from sdpx_gpu import mp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        so = os.path.join(ORACLE_DIR, "liboracle.so")
        src = os.path.join(ORACLE_DIR, "oracle.c")
        if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = ctypes.CDLL(so)
        vp, sz, i32, c = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char
        L.oracle_gemm_exact.argtypes = [i32, c, c, sz, sz, sz, vp, sz, vp, sz, i32, i32, vp, i32]
        L.oracle_gemm_exact.restype = i32
        L.oracle_gemm_entries.argtypes = [i32, c, c, sz, vp, sz, vp, sz, sz, vp, vp, vp, i32]
        L.oracle_gemm_entries.restype = None
        L.oracle_potrf.argtypes = [i32, c, sz, vp, sz]
        L.oracle_potrf.restype = i32
        L.oracle_trsm_lower.argtypes = [i32, sz, vp, vp, sz, i32]
        L.oracle_trsm_lower.restype = None
        L.oracle_potrs.argtypes = [i32, c, sz, sz, vp, sz, vp, sz]
        L.oracle_potrs.restype = None
        for name in ("oracle_div",):
            getattr(L, name).argtypes = [i32, vp, vp, vp]
            getattr(L, name).restype = None
        L.oracle_sqrt.argtypes = [i32, vp, vp]
        L.oracle_sqrt.restype = None
        L.oracle_ldl.argtypes = [i32, vp, sz, sz, vp, sz, vp, vp, vp, i32,
                                 ctypes.POINTER(ctypes.c_longlong), i32]
        L.oracle_ldl.restype = i32
        L.oracle_ldl_mode.argtypes = [i32, vp, sz, sz, vp, sz, vp, vp, vp, i32,
                                      ctypes.POINTER(ctypes.c_longlong), vp, i32]
        L.oracle_ldl_mode.restype = i32
        for name in ("oracle_recip",):
            getattr(L, name).argtypes = [i32, vp, vp]
        for name in ("oracle_sub", "oracle_mul"):
            getattr(L, name).argtypes = [i32, vp, vp, vp]
        _oracle = L
    return _oracle


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb, upper=False, packed=None, threads=8):
    """C oracle: returns None on decline, else the output record array."""
    nl = mp.limbs_of(a if a.size else b)
    if packed is None:
        packed = upper and m == n
    out = mp.zeros(n * (n + 1) // 2 if packed else m * n, nl)
    ok = oracle().oracle_gemm_exact(nl, ta.encode(), tb.encode(), m, n, k, _p(a), lda, _p(b), ldb,
                                    1 if upper else 0, 1 if packed else 0, _p(out), threads)
    return out if ok else None


def oracle_entries(ta, tb, k, a, lda, b, ldb, rows, cols, threads=8):
    nl = mp.limbs_of(a)
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    out = mp.zeros(rows.size, nl)
    oracle().oracle_gemm_entries(nl, ta.encode(), tb.encode(), k, _p(a), lda, _p(b), ldb,
                                 rows.size, _p(rows), _p(cols), _p(out), threads)
    return out


def to_mp(rec):
    """numpy records -> list of mpexact.Mp"""
    n = mp.limbs_of(rec)
    return [mpexact.Mp(n, r["limbs"], r["kind"], r["exponent"]) for r in rec]


def from_mp(vals, n):
    out = mp.zeros(len(vals), n)
    for i, v in enumerate(vals):
        if v is None:
            continue
        out[i]["limbs"] = np.array(v.limbs, dtype=np.uint64)
        out[i]["kind"] = v.kind
        out[i]["exponent"] = v.exponent
    return out


def lcg_operands(m, n, k, spread, seed, nl, ta, tb):
    """The reference's test operands (rns_blas_tests.rs:62-70)."""
    rng = mpexact.Lcg(seed)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    a = mpexact.lcg_random(rng, ar * ac, spread, nl)
    b = mpexact.lcg_random(rng, br * bc, spread, nl)
    return from_mp(a, nl), ar, from_mp(b, nl), br


def oracle_ldl(a, n, signs=None, reg=None, recursive=True, threads=8, mode=None):
    """C oracle LDL^T of the lower triangle of a (n x n records, ld n).
    mode "loaded" (default; recursive=False gives "dots"): factor_loaded,
    ldl_rec from order 256 with the exact-dots retry, else ldl_dots; "rec":
    ldl_rec alone; "dots": ldl_dots.  Returns (ok, l, dinv, regularized)."""
    if mode is None:
        mode = "loaded" if recursive else "dots"
    code = {"dots": 0, "rec": 1, "loaded": 2}[mode]
    nl = mp.limbs_of(a)
    l = a.copy()
    dinv = mp.zeros(n, nl)
    sg = None if signs is None else np.ascontiguousarray(signs, np.int8)
    eps = delta = None
    if reg is not None:
        eps, delta = reg
    cnt = ctypes.c_longlong()
    ok = oracle().oracle_ldl_mode(nl, _p(l), n, n, _p(sg) if sg is not None else None,
                                  0 if sg is None else sg.size, _p(eps), _p(delta), _p(dinv),
                                  code, ctypes.byref(cnt), None, threads)
    return bool(ok), l, dinv, cnt.value


class OraLeaf(ctypes.Structure):
    _fields_ = [("side", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("rinv", ctypes.c_void_p), ("basis", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("coupling", ctypes.c_void_p)]


def oracle_sampled_arrow(problem, nl, border_signs=(-1,), threads=8):
    """C oracle of the sampled Schur build + arrow factor.  Returns
    (ok, border, border_dinv, [(l, dinv, y) per leaf]); ok is None when a
    leaf has 256 or more samples (the recursive leaf path, not restated)."""
    blocks = problem["blocks"]
    t = problem["t"]
    L = oracle()
    L.oracle_sampled_arrow.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.oracle_sampled_arrow.restype = ctypes.c_int
    leaves = (OraLeaf * len(blocks))()
    outs = []
    lp = (ctypes.c_void_p * len(blocks))()
    dp = (ctypes.c_void_p * len(blocks))()
    yp = (ctypes.c_void_p * len(blocks))()
    for i, b in enumerate(blocks):
        leaves[i] = OraLeaf(b["h"], b["k"], _p(b["rinv"]), _p(b["q"]), _p(b["w"]), _p(b["b"]))
        l, d, y = mp.zeros(b["k"] ** 2, nl), mp.zeros(b["k"], nl), mp.zeros(b["k"] * t, nl)
        outs.append((l, d, y))
        lp[i], dp[i], yp[i] = _p(l), _p(d), _p(y)
    border = problem["border"].copy()
    bdinv = mp.zeros(t, nl)
    signs = np.ascontiguousarray(border_signs, np.int8)
    ok = L.oracle_sampled_arrow(nl, len(blocks), leaves, t, _p(border), _p(signs), signs.size,
                                _p(bdinv), lp, dp, yp, threads)
    if ok < 0:  # a leaf of order 256 or more: not restated
        return None, border, bdinv, outs
    return bool(ok), border, bdinv, outs


def oracle_potrf(a, n, lda, uplo="L"):
    """C oracle xpotrf on a copy of the record array; returns (info, a)."""
    nl = mp.limbs_of(a)
    out = a.copy()
    info = oracle().oracle_potrf(nl, uplo.encode(), n, _p(out), lda)
    return info, out


def oracle_trsm_lower(a, n, b, transpose):
    nl = mp.limbs_of(a)
    out = b.copy()
    oracle().oracle_trsm_lower(nl, n, _p(a), _p(out), b.size // max(n, 1), 1 if transpose else 0)
    return out


def oracle_potrs(a, n, lda, b, nrhs, ldb, uplo="L"):
    nl = mp.limbs_of(a)
    out = b.copy()
    oracle().oracle_potrs(nl, uplo.encode(), n, nrhs, _p(a), lda, _p(out), ldb)
    return out


def oracle_sampled_arrow_solve(outs, t, border, bdinv, rhs, nrhs, threads=8):
    """C oracle of ArrowLDLSolver::solve_many on the factors of
    oracle_sampled_arrow (outs: [(l, dinv, y)] per leaf; border with its
    unit factor in the lower triangle): rhs (n x nrhs records) -> x."""
    L = oracle()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.oracle_sampled_arrow_solve.argtypes = [ctypes.c_int, sz, vp, sz, vp, vp, vp, vp, vp, sz, vp, sz,
                                             ctypes.c_int]
    L.oracle_sampled_arrow_solve.restype = ctypes.c_int
    nl = mp.limbs_of(border)
    ks = np.array([d.size for _, d, _ in outs], dtype=np.uint64)
    l_all = np.concatenate([l for l, _, _ in outs]) if outs else mp.zeros(1, nl)
    d_all = np.concatenate([d for _, d, _ in outs]) if outs else mp.zeros(1, nl)
    y_all = np.concatenate([y for _, _, y in outs]) if outs else mp.zeros(1, nl)
    n = int(ks.sum()) + t
    x = np.ascontiguousarray(rhs).reshape(-1).copy()
    L.oracle_sampled_arrow_solve(nl, len(outs), _p(ks), t, _p(l_all), _p(d_all), _p(y_all), _p(border),
                                 _p(bdinv), nrhs, _p(x), n, threads)
    return x


def _leaves_of(problem):
    blocks = problem["blocks"]
    leaves = (OraLeaf * max(1, len(blocks)))()
    for i, b in enumerate(blocks):
        leaves[i] = OraLeaf(b["h"], b["k"], _p(b["rinv"]), _p(b["q"]), _p(b["w"]), _p(b["b"]))
    return leaves


def oracle_sampled_schur(problem, nl, threads=8):
    """C oracle of the sampled Schur blocks: every leaf's H (K x K, lower
    triangle, zeros above) concatenated in leaf order."""
    L = oracle()
    L.oracle_sampled_schur.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int]
    L.oracle_sampled_schur.restype = ctypes.c_int
    blocks = problem["blocks"]
    hs = [mp.zeros(b["k"] ** 2, nl) for b in blocks]
    hp = (ctypes.c_void_p * max(1, len(blocks)))(*[_p(h) for h in hs])
    L.oracle_sampled_schur(nl, len(blocks), _leaves_of(problem), hp, threads)
    return np.concatenate(hs) if hs else mp.zeros(0, nl)


def _kkt_arrays(problem, h_all, nl):
    ks = np.array([b["k"] for b in problem["blocks"]], dtype=np.uint64)
    b_all = np.concatenate([b["b"] for b in problem["blocks"]]) if problem["blocks"] else mp.zeros(1, nl)
    return ks, np.ascontiguousarray(b_all), np.ascontiguousarray(problem["border"])


def oracle_sampled_residual(problem, h_all, rhs, x, threads=8):
    """C oracle of the exact-rows residual RN(b - K x) of one point."""
    L = oracle()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.oracle_sampled_residual.argtypes = [ctypes.c_int, sz, vp, sz, vp, vp, vp, vp, vp, vp, ctypes.c_int]
    L.oracle_sampled_residual.restype = None
    nl = mp.limbs_of(rhs)
    ks, b_all, cmat = _kkt_arrays(problem, h_all, nl)
    e = mp.zeros(rhs.size, nl)
    L.oracle_sampled_residual(nl, len(problem["blocks"]), _p(ks), problem["t"], _p(h_all), _p(b_all), _p(cmat),
                              _p(np.ascontiguousarray(rhs)), _p(np.ascontiguousarray(x)), _p(e), threads)
    return e


def oracle_sampled_refine(problem, h_all, outs, border, bdinv, rhs, nrhs, reltol, abstol, stop_ratio, max_iter,
                          threads=8):
    """C oracle of DirectLDLKKTSolver::solve_many with stationary refinement
    on the sampled arrow KKT: returns (x, ok per column, norms, refinements)."""
    L = oracle()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.oracle_sampled_refine.argtypes = [ctypes.c_int, sz, vp, sz, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                        ctypes.c_uint, sz, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.c_int]
    L.oracle_sampled_refine.restype = ctypes.c_int
    nl = mp.limbs_of(border)
    ks, b_all, cmat = _kkt_arrays(problem, h_all, nl)
    l_all = np.concatenate([l for l, _, _ in outs]) if outs else mp.zeros(1, nl)
    d_all = np.concatenate([d for _, d, _ in outs]) if outs else mp.zeros(1, nl)
    y_all = np.concatenate([y for _, _, y in outs]) if outs else mp.zeros(1, nl)

    def rec(v):
        return v if isinstance(v, np.ndarray) else mp.from_float64(np.array([float(v)]), nl)

    rt, at, sr = rec(reltol), rec(abstol), rec(stop_ratio)
    x = mp.zeros(rhs.size, nl)
    ok = np.zeros(max(1, nrhs), np.int32)
    norms = mp.zeros(max(1, nrhs), nl)
    cnt = ctypes.c_longlong()
    L.oracle_sampled_refine(nl, len(outs), _p(ks), problem["t"], _p(h_all), _p(b_all), _p(cmat), _p(l_all),
                            _p(d_all), _p(y_all), _p(border), _p(bdinv), _p(rt), _p(at), _p(sr), max_iter, nrhs,
                            _p(np.ascontiguousarray(rhs)), _p(x), _p(ok), _p(norms), ctypes.byref(cnt), threads)
    return x, [bool(v) for v in ok[:nrhs]], norms[:nrhs], cnt.value


def oracle_gemm_rns(ta, tb, m, n, k, a, lda, b, ldb, upper=False, packed=None, threads=8):
    """The CPU residue product (the reference's algorithm, the CPU
    baseline's exact product): None on decline, else the output records."""
    L = oracle()
    L.oracle_gemm_rns.argtypes = L.oracle_gemm_exact.argtypes
    L.oracle_gemm_rns.restype = ctypes.c_int
    nl = mp.limbs_of(a if a.size else b)
    if packed is None:
        packed = upper and m == n
    out = mp.zeros(n * (n + 1) // 2 if packed else m * n, nl)
    ok = L.oracle_gemm_rns(nl, ta.encode(), tb.encode(), m, n, k, _p(a), lda, _p(b), ldb,
                           1 if upper else 0, 1 if packed else 0, _p(out), threads)
    return out if ok else None


def oracle_set_rns_products(on):
    """Make the restatements' exact products the CPU residue product where
    the reference takes it (CPU baseline timing); off: exact dots."""
    oracle().oracle_set_rns_products.argtypes = [ctypes.c_int]
    oracle().oracle_set_rns_products(1 if on else 0)
