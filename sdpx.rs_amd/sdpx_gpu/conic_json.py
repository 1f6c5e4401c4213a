"""Reader of the solver's conic JSON input (JsonProblem,
crates/solver/src/solver/default/json.rs:15-36: P and A as CscMatrix fields
{m, n, colptr, rowval, nzval}, q, b, cones as serde's externally tagged
SupportedConeT variants, optional settings and sampled blocks), with the
reference's input checks (check_dimensions, default/solver.rs:346-438, and
its messages), and its mapping onto this library's condensed Float64
assembly (PSD triangle and nonnegative cones).

Values are JSON numbers or decimal strings.  With nl=None they are read as
binary64; with nl limbs, decimal strings are rounded once to nearest-even at
64 nl bits (sdpb_json.decimal_record) and numbers converted exactly.
Cones come back as tuples: ("zero", n), ("nonneg", n), ("soc", n), ("exp",),
("pow", a), ("genpow", alphas, d), ("psd", n), ("oprel", n),
("oprel_limit", n, m).
"""
import json
from fractions import Fraction

import numpy as np

from . import mp
from .sdpb_json import decimal_record


class ConicInputError(ValueError):
    pass


_TAGS = {"ZeroConeT": "zero", "NonnegativeConeT": "nonneg", "SecondOrderConeT": "soc",
         "ExponentialConeT": "exp", "PowerConeT": "pow", "GenPowerConeT": "genpow", "PSDTriangleConeT": "psd",
         "OpRelEntropyConeT": "oprel", "OpRelEntropyLimitConeT": "oprel_limit"}


def _t(n):
    return n * (n + 1) // 2


def _value(v, nl):
    if isinstance(v, bool) or not isinstance(v, (int, float, str)):
        raise ConicInputError("values must be numbers or decimal strings")
    if nl is None:
        return float(v)
    if isinstance(v, str):
        return decimal_record(v, nl)
    return _exact_record(Fraction(v), nl)


def _exact_record(x, nl):
    """A binary64 or integer value (exact as a fraction) as a record rounded
    once to 64 nl bits."""
    if x == 0:
        return (0,) * nl, mp.ZERO_KIND, 0
    neg = x < 0
    x = abs(x)
    P = 64 * nl
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    e += 1  # x in [2^(e-1), 2^e)
    scaled = x * Fraction(2) ** (P - e)
    q, r = divmod(scaled.numerator, scaled.denominator)
    if 2 * r > scaled.denominator or (2 * r == scaled.denominator and q & 1):
        q += 1
    if q >> P:
        q >>= 1
        e += 1
    limbs = tuple((q >> (64 * i)) & ((1 << 64) - 1) for i in range(nl))
    return limbs, -mp.REGULAR_KIND if neg else mp.REGULAR_KIND, e


def _values(arr, nl, what):
    if not isinstance(arr, list):
        raise ConicInputError(f"{what} must be an array")
    vals = [_value(v, nl) for v in arr]
    if nl is None:
        return np.array(vals, dtype=np.float64)
    out = mp.zeros(len(vals), nl)
    for i, (limbs, kind, e) in enumerate(vals):
        out["limbs"][i] = limbs
        out["kind"][i] = kind
        out["exponent"][i] = e
    return out


def _count(v, what):
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise ConicInputError(f"{what} must be a nonnegative integer")
    return v


def _csc(obj, nl, what):
    if not isinstance(obj, dict):
        raise ConicInputError(f"{what} must be an object")
    for f in ("m", "n", "colptr", "rowval", "nzval"):
        if f not in obj:
            raise ConicInputError(f"missing field `{f}` in {what}")
    m, n = _count(obj["m"], f"{what}.m"), _count(obj["n"], f"{what}.n")
    colptr = [_count(v, f"{what}.colptr") for v in obj["colptr"]]
    rowval = [_count(v, f"{what}.rowval") for v in obj["rowval"]]
    nzval = _values(obj["nzval"], nl, f"{what}.nzval")
    nnz = len(rowval)
    if len(colptr) != n + 1 or colptr[0] != 0 or colptr[-1] != nnz or len(nzval) != nnz or \
            any(colptr[j] > colptr[j + 1] for j in range(n)) or any(r >= m for r in rowval):
        raise ConicInputError(f"{what} is not a valid CSC matrix")
    return {"m": m, "n": n, "colptr": np.array(colptr, dtype=np.int64),
            "rowval": np.array(rowval, dtype=np.int64), "nzval": nzval}


def _cone(obj):
    if isinstance(obj, str) and obj in _TAGS:  # a unit variant written as its name
        obj = {obj: []}
    if not isinstance(obj, dict) or len(obj) != 1:
        raise ConicInputError("a cone must be an object with one variant")
    (tag, arg), = obj.items()
    if tag not in _TAGS:
        raise ConicInputError(f"unknown cone variant `{tag}`")
    kind = _TAGS[tag]
    if kind == "exp":
        if arg not in ([], None):
            raise ConicInputError("ExponentialConeT takes no parameters")
        return ("exp",)
    if kind == "pow":
        return ("pow", float(_value(arg, None)))
    if kind == "genpow":
        if not isinstance(arg, list) or len(arg) != 2:
            raise ConicInputError("GenPowerConeT takes (alphas, dim)")
        return ("genpow", [float(_value(a, None)) for a in arg[0]], _count(arg[1], "GenPowerConeT dim"))
    if kind == "oprel_limit":
        if not isinstance(arg, list) or len(arg) != 2:
            raise ConicInputError("OpRelEntropyLimitConeT takes (n, m)")
        return ("oprel_limit", _count(arg[0], "n"), _count(arg[1], "m"))
    return (kind, _count(arg, tag))


def cone_size(cone):
    """Rows of a cone, or None for invalid parameters (check_dimensions'
    "invalid cone dimensions or parameters")."""
    kind = cone[0]
    if kind in ("zero", "nonneg", "soc"):
        return cone[1]
    if kind == "exp":
        return 3
    if kind == "pow":
        return 3 if 0.0 < cone[1] < 1.0 else None
    if kind == "genpow":
        alphas = cone[1]
        total = 0.0
        for a in alphas:  # the reference's sequential sum
            total += a
        if not all(a > 0 for a in alphas) or not abs(1.0 - total) < 2.0 ** -52 * len(alphas) * 0.5:
            return None
        return len(alphas) + cone[2]
    if kind == "psd":
        return _t(cone[1])
    if kind == "oprel":
        return 3 * _t(cone[1]) if cone[1] > 0 else None
    n, m = cone[1], cone[2]
    k = n - m
    return _t(n) + m * k + 2 * _t(k) if k > 0 else None


def check_dimensions(P, q, A, b, cones):
    """The reference's check_dimensions (default/solver.rs:346-438), its
    messages and their order."""
    m, n = len(b), len(q)
    bf = b if b.dtype == np.float64 else mp.to_float64(b)
    p = 0
    for cone in cones:
        size = cone_size(cone)
        if size is None:
            raise ConicInputError("invalid cone dimensions or parameters")
        rows = bf[p:p + size]
        ok = np.isfinite(rows) | ((cone[0] == "nonneg") & (rows == np.inf))
        if not np.all(ok):
            raise ConicInputError("b must be finite, except +inf in NonnegativeConeT rows")
        p += size
    if m != A["m"]:
        raise ConicInputError("A and b incompatible dimensions")
    if p != m:
        raise ConicInputError("Constraint dimensions inconsistent with size of cones")
    if n != A["n"]:
        raise ConicInputError("A and q incompatible dimensions")
    if n != P["n"]:
        raise ConicInputError("P and q incompatible dimensions")
    if P["m"] != P["n"]:
        raise ConicInputError("P not square")


def read_conic_json(src, nl=None):
    """Read a conic JSON problem from a path, a file object or a string of
    JSON text; returns {"P", "q", "A", "b", "cones", "settings", "sampled"}."""
    if hasattr(src, "read"):
        obj = json.load(src)
    elif isinstance(src, str) and src.lstrip().startswith("{"):
        obj = json.loads(src)
    else:
        with open(src) as f:
            obj = json.load(f)
    if not isinstance(obj, dict):
        raise ConicInputError("the input must be a JSON object")
    for f in ("P", "q", "A", "b", "cones"):
        if f not in obj:
            raise ConicInputError(f"missing field `{f}`")
    P = _csc(obj["P"], nl, "P")
    A = _csc(obj["A"], nl, "A")
    q = _values(obj["q"], nl, "q")
    b = _values(obj["b"], nl, "b")
    if not isinstance(obj["cones"], list):
        raise ConicInputError("cones must be an array")
    cones = [_cone(c) for c in obj["cones"]]
    check_dimensions(P, q, A, b, cones)
    sampled = []
    for s in obj.get("sampled", []) or []:
        fields = ("row_start", "column_start", "dim", "basis_rows", "basis_cols", "basis", "weights")
        if not isinstance(s, dict) or any(f not in s for f in fields):
            raise ConicInputError("a sampled block needs " + ", ".join(fields))
        blk = {f: _count(s[f], f) for f in fields[:5]}
        blk["basis"] = _values(s["basis"], nl, "basis")
        blk["weights"] = _values(s["weights"], nl, "weights")
        if len(blk["basis"]) != blk["basis_rows"] * blk["basis_cols"]:
            raise ConicInputError("sampled basis size differs from basis_rows x basis_cols")
        if len(blk["weights"]) != _t(blk["dim"]) * blk["basis_cols"]:
            raise ConicInputError("sampled weights size differs from dim (dim + 1) / 2 x basis_cols")
        sampled.append(blk)
    return {"P": P, "q": q, "A": A, "b": b, "cones": cones, "settings": obj.get("settings", {}) or {},
            "sampled": sampled}


def _cone_json(cone):
    kind = cone[0]
    tag = {v: k for k, v in _TAGS.items()}[kind]
    if kind == "exp":
        return {tag: []}
    if kind in ("genpow", "oprel_limit"):
        return {tag: [cone[1], cone[2]]}
    return {tag: cone[1]}


def write_conic_json(problem):
    """The JSON text of a binary64 problem in the reader's form (P, A as
    CSC objects, q, b, cones, settings)."""
    csc = lambda M: {"m": int(M["m"]), "n": int(M["n"]), "colptr": [int(v) for v in M["colptr"]],
                     "rowval": [int(v) for v in M["rowval"]], "nzval": [float(v) for v in M["nzval"]]}
    return json.dumps({"P": csc(problem["P"]), "q": [float(v) for v in problem["q"]], "A": csc(problem["A"]),
                       "b": [float(v) for v in problem["b"]], "cones": [_cone_json(c) for c in problem["cones"]],
                       "settings": problem.get("settings", {})})


def dense(M):
    """A CSC field as a dense binary64 array."""
    out = np.zeros((M["m"], M["n"]))
    vals = M["nzval"] if M["nzval"].dtype == np.float64 else mp.to_float64(M["nzval"])
    for j in range(M["n"]):
        for p in range(M["colptr"][j], M["colptr"][j + 1]):
            out[M["rowval"][p], j] = vals[p]
    return out


def condensed_inputs(problem, rinvs, weights):
    """The condensed Float64 assembly's inputs for a problem of PSD triangle,
    nonnegative and zero cones: P's upper CSC, one PSD block per PSD cone
    (rinvs[i] its NT factor; a_svec its rows of A over the columns that
    touch it, those columns), the nonnegative rows of A with their weights
    (H^-1's diagonal), and the zero cones' row ranges (retained rows, left
    to the caller).  Other cones raise ConicInputError."""
    if problem["sampled"]:
        raise ConicInputError("sampled blocks go through the arrow KKT (sdpb_json.arrow_problem)")
    P, A = dense(problem["P"]), dense(problem["A"])
    Pu = np.triu(P)
    colptr, rowval, nzval = [0], [], []
    for j in range(Pu.shape[1]):
        for i in range(j + 1):
            if Pu[i, j] != 0:
                rowval.append(i)
                nzval.append(Pu[i, j])
        colptr.append(len(rowval))
    blocks, orth_rows, retained = [], [], []
    p = 0
    psd = 0
    for cone in problem["cones"]:
        size = cone_size(cone)
        if cone[0] == "psd":
            rows = A[p:p + size]
            cols = np.nonzero(np.any(rows != 0, axis=0))[0]
            blocks.append((rinvs[psd], rows[:, cols], cols))
            psd += 1
        elif cone[0] == "nonneg":
            orth_rows.extend(range(p, p + size))
        elif cone[0] == "zero":
            retained.append((p, p + size))
        else:
            raise ConicInputError(f"the condensed Float64 assembly takes PSD, nonnegative and zero cones, not {cone[0]}")
        p += size
    B = A[orth_rows] if orth_rows else None
    w = np.asarray(weights, dtype=np.float64) if orth_rows else None
    if orth_rows and len(w) != len(orth_rows):
        raise ConicInputError("one weight per nonnegative row")
    return (colptr, rowval, nzval), blocks, B, w, retained
