"""Reader of SDPB's uncompressed `pmp2sdp --outputFormat=json` directories
(the sampled input of crates/solver/src/solver/sampled/json.rs:163-313,
read_sdpb_sampled), at the precision of the records it returns.

A directory holds control.json ({"num_blocks", optional "command"}),
objectives.json ({"constant", "b"}), and per block i block_info_i.json
({"dim", "num_points"}) and block_data_i.json ({"bilinear_bases_even",
"bilinear_bases_odd", "c", "B"}); normalization.json and pmp_info.json may
be present and are not read.  Every coefficient is a decimal string, parsed
exactly and rounded once to nearest-even at 64 * limbs bits (as MPFR's
string conversion at the working precision; never through binary64).

The result mirrors SampledJsonProblem (json.rs:23-35):
  q (n records), A in CSC (the B rows as columns, zeros dropped), b (the
  equality right-hand sides, then zeros for the PSD rows), cones
  ([("zero", E)] then ("psd", side) per non-empty parity basis),
  sampled blocks (row_start, column_start, dim, basis_rows, basis_cols,
  basis: column-major basis_rows x num_points, weights: -1 each), grams
  (block_index, parity, row_start, side), objective_constant,
  num_equalities.
Malformed input raises SampledInputError with the reference's messages.
"""
import json
import os
import re
from fractions import Fraction

import numpy as np

from . import mp

_DECIMAL = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")
_EMIN, _EMAX = 1 - (1 << 30), (1 << 30) - 1


class SampledInputError(ValueError):
    pass


def _bad(message):
    return SampledInputError(message)


def decimal_record(text, nl):
    """A decimal string -> (limbs tuple, kind, exponent) of its value rounded
    once to nearest-even at 64 nl bits."""
    if not isinstance(text, str):
        raise _bad("sampled coefficients must be finite decimal strings")
    if not text or not all(c.isdigit() or c in "+-.eE" for c in text):
        raise _bad("sampled coefficients must be finite decimal strings")
    m = _DECIMAL.match(text)
    if not m:
        raise _bad("invalid decimal coefficient")
    negative = text.startswith("-")
    mantissa, exp_part = m.group(1), m.group(2)
    digits_nonzero = any(c in "123456789" for c in mantissa)
    if not digits_nonzero:
        return (0,) * nl, -2 if negative else 2, 0
    e10 = int(exp_part[1:]) if exp_part else 0
    if abs(e10) > 400_000_000:  # beyond the exponent range at any mantissa
        if e10 > 0:
            raise _bad("sampled coefficients must be finite decimal strings")
        raise _bad("sampled coefficient underflow")
    a = abs(Fraction(text))
    p = 64 * nl
    e = a.numerator.bit_length() - a.denominator.bit_length()  # 2^e <= a < 2^(e+2)
    if Fraction(2) ** e > a:
        e -= 1
    if Fraction(2) ** (e + 1) <= a:
        e += 1
    scaled = a * Fraction(2) ** (p - 1 - e)  # in [2^(p-1), 2^p)
    q, r = divmod(scaled.numerator, scaled.denominator)
    if 2 * r > scaled.denominator or (2 * r == scaled.denominator and q & 1):
        q += 1
    if q == 1 << p:
        q >>= 1
        e += 1
    exponent = e + 1
    if exponent > _EMAX:
        raise _bad("sampled coefficients must be finite decimal strings")
    if exponent < _EMIN:
        raise _bad("sampled coefficient underflow")
    limbs = tuple((q >> (64 * i)) & ((1 << 64) - 1) for i in range(nl))
    return limbs, -3 if negative else 3, exponent


def _records(texts, nl):
    out = mp.zeros(len(texts), nl)
    for i, t in enumerate(texts):
        limbs, kind, exponent = decimal_record(t, nl)
        out[i]["limbs"] = np.array(limbs, dtype=np.uint64)
        out[i]["kind"] = kind
        out[i]["exponent"] = exponent
    return out


def _read(directory, name):
    with open(os.path.join(directory, name)) as f:
        return json.load(f)


def _fields(obj, required, optional=(), what=""):
    if not isinstance(obj, dict):
        raise _bad(f"{what}: expected an object")
    extra = set(obj) - set(required) - set(optional)
    if extra:
        raise _bad(f"{what}: unknown field {sorted(extra)[0]}")
    missing = [k for k in required if k not in obj]
    if missing:
        raise _bad(f"{what}: missing field {missing[0]}")


def _count(v, what):
    if not isinstance(v, int) or isinstance(v, bool) or v < 0:
        raise _bad(f"{what}: expected a non-negative integer")
    return v


def _triangle(n):
    return n * (n + 1) // 2


def read_sdpb_sampled(directory, nl):
    """Read one sampled directory at 64 nl bits (see the module comment)."""
    control = _read(directory, "control.json")
    _fields(control, ("num_blocks",), ("command",), "control.json")
    num_blocks = _count(control["num_blocks"], "num_blocks")
    for name in os.listdir(directory):
        if not name.endswith(".json"):
            continue
        known = name in ("control.json", "objectives.json", "normalization.json", "pmp_info.json")
        for prefix in ("block_info_", "block_data_"):
            s = name[len(prefix):-5] if name.startswith(prefix) else None
            if s is not None and s.isdigit() and str(int(s)) == s and int(s) < num_blocks:
                known = True
        if not known:
            raise _bad(f"unexpected sampled JSON file: {name}")
    objectives = _read(directory, "objectives.json")
    _fields(objectives, ("constant", "b"), (), "objectives.json")
    constant = _records([objectives["constant"]], nl)
    rhs = list(objectives["b"]) if isinstance(objectives["b"], list) else None
    if rhs is None:
        raise _bad("objectives.json: b must be an array")
    equalities = len(rhs)
    cones = [("zero", equalities)] if equalities else []
    q, colptr, rowval, nzval, sampled, grams = [], [0], [], [], [], []
    rows = equalities
    for bi in range(num_blocks):
        info = _read(directory, f"block_info_{bi}.json")
        data = _read(directory, f"block_data_{bi}.json")
        _fields(info, ("dim", "num_points"), (), f"block_info_{bi}.json")
        _fields(data, ("bilinear_bases_even", "bilinear_bases_odd", "c", "B"), (), f"block_data_{bi}.json")
        dim, points = _count(info["dim"], "dim"), _count(info["num_points"], "num_points")
        count = _triangle(dim) * points
        if dim == 0 or points == 0 or len(data["c"]) != count or len(data["B"]) != count:
            raise _bad("inconsistent sampled block dimensions")
        column_start = len(q)
        for c, brow in zip(data["c"], data["B"]):
            if not isinstance(brow, list) or len(brow) != equalities:
                raise _bad("sampled B width mismatch")
            q.append(c)
            recs = _records(brow, nl) if brow else mp.zeros(0, nl)
            for r, rec in enumerate(recs):
                if abs(int(rec["kind"])) != 2:
                    rowval.append(r)
                    nzval.append(rec)
            colptr.append(len(nzval))
        for parity, key in enumerate(("bilinear_bases_even", "bilinear_bases_odd")):
            basis = data[key]
            if not isinstance(basis, list):
                raise _bad("sampled basis must be an array")
            if not basis:
                continue
            height = len(basis)
            if any(not isinstance(r, list) or len(r) != points for r in basis):
                raise _bad("sampled basis width mismatch")
            side = dim * height
            values = _records([basis[row][k] for k in range(points) for row in range(height)], nl)
            weights = mp.from_float64(np.full(count, -1.0), nl)
            sampled.append({"row_start": rows, "column_start": column_start, "dim": dim, "basis_rows": height,
                            "basis_cols": points, "basis": values, "weights": weights})
            grams.append({"block_index": bi, "parity": parity, "row_start": rows, "side": side})
            cones.append(("psd", side))
            rows += _triangle(side)
    n = len(q)
    b = mp.zeros(rows, nl)
    if equalities:
        b[:equalities] = _records(rhs, nl)
    return {
        "objective_constant": constant[0], "num_equalities": equalities,
        "q": _records(q, nl) if q else mp.zeros(0, nl),
        "A": {"m": rows, "n": n, "colptr": np.array(colptr, np.int64), "rowval": np.array(rowval, np.int64),
              "nzval": np.array(nzval, dtype=mp.dtype(nl)) if nzval else mp.zeros(0, nl)},
        "b": b, "cones": cones, "sampled": sampled, "grams": grams,
    }


def arrow_problem(parsed, nl, rinv=None, weights=None, border=None):
    """The sampled arrow KKT of an interior-point iteration on a read
    problem whose blocks are scalar (dim 1) with one non-empty basis parity:
    one leaf per sampled block (its samples are the leaf's rows), coupled to
    the equality rows by the B rows of its columns; rinv[i] (h x h records,
    default identity) and weights[i] (default the blocks' -1s) stand for the
    iterate; border: the t x t border block (default zeros).  Returns the
    problem dictionary of workloads.sampled_problem."""
    t = parsed["num_equalities"]
    A = parsed["A"]
    seen = set()
    blocks = []
    for i, blk in enumerate(parsed["sampled"]):
        if blk["dim"] != 1:
            raise _bad("the arrow form takes scalar (dim 1) sampled blocks")
        if blk["column_start"] in seen:
            raise _bad("the arrow form takes one basis parity per block")
        seen.add(blk["column_start"])
        h, k = blk["basis_rows"], blk["basis_cols"]
        r = rinv[i] if rinv is not None else mp.from_float64(np.eye(h).reshape(-1), nl)
        w = weights[i] if weights is not None else blk["weights"]
        coupling = mp.zeros(k * t, nl)
        for j in range(k):
            col = blk["column_start"] + j
            for p in range(int(A["colptr"][col]), int(A["colptr"][col + 1])):
                coupling[j + int(A["rowval"][p]) * k] = A["nzval"][p]
        blocks.append({"h": h, "k": k, "rinv": r, "q": blk["basis"], "w": w, "b": coupling})
    return {"blocks": blocks, "t": t, "border": border if border is not None else mp.zeros(t * t, nl)}
