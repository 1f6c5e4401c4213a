"""Conic JSON input (sdpx_gpu/conic_json.py): the reference's basic SDP
(crates/solver/tests/it/basic_sdp.rs:5-39) as a JSON fixture, its known
solution through the svec convention the condensed assembly uses, a round
trip, decimal strings rounded once at 256 bits, the reference's input
checks and messages (check_dimensions, default/solver.rs:346-438), and the
mapping onto the condensed assembly's inputs."""
import json
import os
import re
from fractions import Fraction

import numpy as np
import pytest

from sdpx_gpu import conic_json as cj
from sdpx_gpu import ipm, mp

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "basic_sdp.json")
# basic_sdp_solution (basic_sdp.rs:28-39)
REFSOL = [-3.0729833267361095, 0.3696004167288786, -0.022226685581313674, 0.31441213129613066,
          -0.026739700851545107, -0.016084530571308823]
REFOBJ = 4.840076866013861


def test_basic_sdp_fixture_and_known_solution():
    """min 1/2 |x|^2 s.t. b - x in the PSD triangle cone (P = A = I): x* =
    b - proj(b) = svec(V min(L, 0) V^T) for smat(b) = V L V^T.  Agreeing with
    the reference's solution pins the triangle's order and sqrt(2) scaling
    that the condensed assembly's svec uses."""
    pr = cj.read_conic_json(FIXTURE)
    assert pr["cones"] == [("psd", 3)]
    assert np.array_equal(cj.dense(pr["P"]), np.eye(6)) and np.array_equal(cj.dense(pr["A"]), np.eye(6))
    assert np.array_equal(pr["b"], [-3.0, 1.0, 4.0, 1.0, 2.0, 5.0]) and not pr["q"].any()
    lam, V = np.linalg.eigh(ipm.smat(pr["b"], 3))
    x = ipm.svec((V * np.minimum(lam, 0.0)) @ V.T)
    assert np.abs(x - REFSOL).max() <= 1e-9
    assert abs(0.5 * x @ x - REFOBJ) <= 1e-9
    # the other orders a reader could take miss it
    xr = ipm.svec(((V * np.minimum(lam, 0.0)) @ V.T).T)
    bad = ipm.smat(np.array(pr["b"])[[0, 1, 3, 2, 4, 5]], 3)
    lam2, V2 = np.linalg.eigh(bad)
    assert np.abs(ipm.svec((V2 * np.minimum(lam2, 0.0)) @ V2.T) - REFSOL).max() > 1e-3
    assert np.allclose(xr, x)  # symmetric: transposition changes nothing


def test_round_trip_and_cone_variants():
    rng = np.random.default_rng(3)
    cones = [("zero", 2), ("nonneg", 3), ("soc", 4), ("exp",), ("pow", 0.3), ("genpow", [0.25, 0.75], 2),
             ("psd", 3), ("oprel", 2), ("oprel_limit", 3, 1)]
    m = sum(cj.cone_size(c) for c in cones)
    n = 5
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)

    def csc(M):
        cp, rv, nz = [0], [], []
        for j in range(M.shape[1]):
            for i in range(M.shape[0]):
                if M[i, j] != 0:
                    rv.append(i)
                    nz.append(M[i, j])
            cp.append(len(rv))
        return {"m": M.shape[0], "n": M.shape[1], "colptr": np.array(cp), "rowval": np.array(rv),
                "nzval": np.array(nz)}

    P = np.triu(rng.standard_normal((n, n)))
    pr = {"P": csc(P), "q": rng.standard_normal(n), "A": csc(A), "b": rng.standard_normal(m), "cones": cones,
          "settings": {"max_iter": 50}}
    back = cj.read_conic_json(cj.write_conic_json(pr))
    assert back["cones"] == cones and back["settings"] == {"max_iter": 50}
    assert np.array_equal(cj.dense(back["A"]), A) and np.array_equal(cj.dense(back["P"]), P)
    assert np.array_equal(back["q"], pr["q"]) and np.array_equal(back["b"], pr["b"])


def test_decimal_strings_round_once_at_256_bits():
    doc = json.load(open(FIXTURE))
    doc["b"] = ["0.1", "-1e-30", "3", "1.5", "2", "5"]
    doc["A"]["nzval"] = [1, 1, 1, 1, 1, "0.333333333333333333333333333333333333"]
    pr = cj.read_conic_json(json.dumps(doc), nl=4)
    P = 256
    for rec, text in zip(pr["b"], doc["b"]):
        x = Fraction(text)
        e = 0
        while Fraction(2) ** e <= abs(x):
            e += 1
        while Fraction(2) ** (e - 1) > abs(x):
            e -= 1
        scaled = abs(x) * Fraction(2) ** (P - e)
        q, r = divmod(scaled.numerator, scaled.denominator)
        if 2 * r > scaled.denominator or (2 * r == scaled.denominator and q & 1):
            q += 1
        got = sum(int(v) << (64 * i) for i, v in enumerate(rec["limbs"]))
        assert got == q and int(rec["exponent"]) == e and (int(rec["kind"]) < 0) == (x < 0)
    assert np.array_equal(mp.to_float64(pr["A"]["nzval"])[:5], np.ones(5))


@pytest.mark.parametrize("edit, message", [
    (lambda d: d.update(b=d["b"][:5]), "A and b incompatible dimensions"),
    (lambda d: d.update(cones=[{"PSDTriangleConeT": 2}, {"NonnegativeConeT": 2}]),
     "Constraint dimensions inconsistent with size of cones"),
    (lambda d: d.update(q=d["q"][:5]), "A and q incompatible dimensions"),
    (lambda d: d["P"].update(m=6, n=5, colptr=d["P"]["colptr"][:6], rowval=d["P"]["rowval"][:5],
                             nzval=d["P"]["nzval"][:5]), "P and q incompatible dimensions"),
    (lambda d: d.update(cones=[{"PowerConeT": 1.5}, {"NonnegativeConeT": 3}]),
     "invalid cone dimensions or parameters"),
    (lambda d: d.update(b=["nan", 1, 4, 1, 2, 5]), "b must be finite, except +inf in NonnegativeConeT rows"),
])
def test_reference_input_checks(edit, message):
    doc = json.load(open(FIXTURE))
    edit(doc)
    text = json.dumps(doc).replace('"nan"', "NaN")
    with pytest.raises(cj.ConicInputError, match=re.escape(message)):
        cj.read_conic_json(text)


def test_infinite_bound_allowed_in_orthant_rows_only():
    doc = json.load(open(FIXTURE))
    doc["cones"] = [{"NonnegativeConeT": 6}]
    doc["b"][2] = 1e400  # +inf
    cj.read_conic_json(json.dumps(doc).replace("1e+400", "Infinity").replace("inf", "Infinity"))
    doc["cones"] = [{"ZeroConeT": 6}]
    with pytest.raises(cj.ConicInputError, match="b must be finite"):
        cj.read_conic_json(json.dumps(doc).replace("inf", "Infinity"))


def test_condensed_inputs_of_a_mixed_problem():
    """PSD, nonnegative and zero cones onto the condensed assembly's inputs:
    the Schur matrix they define, P + sum A_b^T H_b A_b + B^T diag(w) B,
    equals P + A_K^T H A_K formed densely from the problem's own A."""
    rng = np.random.default_rng(5)
    cones = [("zero", 2), ("psd", 3), ("nonneg", 4), ("psd", 2)]
    m = sum(cj.cone_size(c) for c in cones)
    n = 7
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.6)
    Pd = rng.standard_normal((n, n))
    Pd = Pd @ Pd.T

    def csc(M):
        cp, rv, nz = [0], [], []
        for j in range(M.shape[1]):
            for i in range(M.shape[0]):
                if M[i, j] != 0:
                    rv.append(i)
                    nz.append(M[i, j])
            cp.append(len(rv))
        return {"m": M.shape[0], "n": M.shape[1], "colptr": cp, "rowval": rv, "nzval": nz}

    doc = {"P": csc(np.triu(Pd)), "q": [0.0] * n, "A": csc(A), "b": [0.0] * m,
           "cones": [cj._cone_json(c) for c in cones]}
    pr = cj.read_conic_json(json.dumps(doc))
    rinvs = [np.eye(3) + 0.1 * rng.standard_normal((3, 3)), np.eye(2) + 0.1 * rng.standard_normal((2, 2))]
    w = np.exp(rng.standard_normal(4))
    (cp, rv, nz), blocks, B, ww, retained = cj.condensed_inputs(pr, rinvs, w)
    assert retained == [(0, 2)]
    S = np.zeros((n, n))
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            S[rv[p], j] += nz[p]
            if rv[p] != j:
                S[j, rv[p]] += nz[p]
    for rinv, a_svec, cols in blocks:
        W = rinv.T @ rinv
        k = rinv.shape[0]
        H = np.stack([ipm.svec(W @ ipm.smat(e, k) @ W) for e in np.eye(k * (k + 1) // 2)], axis=1)
        S[np.ix_(cols, cols)] += a_svec.T @ H @ a_svec
    S += B.T @ (ww[:, None] * B)
    want = Pd.copy()
    row, psd_i = 0, 0
    for cone in cones:
        size = cj.cone_size(cone)
        Ak = A[row:row + size]
        if cone[0] == "psd":
            rinv = rinvs[psd_i]
            psd_i += 1
            k = rinv.shape[0]
            W = rinv.T @ rinv
            H = np.stack([ipm.svec(W @ ipm.smat(e, k) @ W) for e in np.eye(size)], axis=1)
            want += Ak.T @ H @ Ak
        elif cone[0] == "nonneg":
            want += Ak.T @ (w[:, None] * Ak)
        row += size
    assert np.abs(S - want).max() <= 1e-12 * np.abs(want).max()
    with pytest.raises(cj.ConicInputError, match="not soc"):
        doc2 = dict(doc, cones=[cj._cone_json(c) for c in [("zero", 2), ("soc", 6), ("nonneg", 4), ("psd", 2)]])
        cj.condensed_inputs(cj.read_conic_json(json.dumps(doc2)), rinvs, w)
