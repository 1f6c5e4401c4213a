"""Reader of SDPB pmp2sdp sampled JSON directories (sampled/json.rs:163-313):
the reference's own fixture and known answers (crates/solver/tests/it/
native_input.rs:160-236), exact decimal rounding against rational
arithmetic (ties to even), the reference's input checks, and a scalar
sampled directory read back into the arrow KKT the factor takes (CPU)."""
import json
from fractions import Fraction

import numpy as np
import pytest

import pins
from sdpx_gpu import mp, sdpb_json, workloads

# the reference's fixture (native_input.rs:160-172)
FIXTURE = {
    "control.json": {"num_blocks": 1, "command": "pmp2sdp"},
    "objectives.json": {"constant": "3", "b": ["1"]},
    "block_info_0.json": {"dim": 2, "num_points": 1},
    "block_data_0.json": {"bilinear_bases_even": [["2"]], "bilinear_bases_odd": [], "c": ["1", "0", "1"],
                          "B": [["1"], ["0"], ["1"]]},
}


def write(path, files):
    for name, obj in files.items():
        (path / name).write_text(obj if isinstance(obj, str) else json.dumps(obj))


def apply_operator(parsed, x):
    """A x of the read problem in rationals: the B rows, then per sampled
    block X = sum_p x_p w_p sym(E_rs) (x) v_k v_k^T over its columns p in
    (s, r <= s, k) order, k fastest, written to its rows as svec (upper
    triangle column by column, off-diagonals times sqrt 2, kept as a
    coefficient of sqrt 2)."""
    A = parsed["A"]
    out = [Fraction(0)] * A["m"]
    sq = [Fraction(0)] * A["m"]  # coefficients of sqrt(2)
    vals = pins.dec(A["nzval"]) if A["nzval"].size else []
    for col in range(A["n"]):
        for p in range(int(A["colptr"][col]), int(A["colptr"][col + 1])):
            out[int(A["rowval"][p])] += vals[p] * x[col]
    for blk in parsed["sampled"]:
        dim, h, npts = blk["dim"], blk["basis_rows"], blk["basis_cols"]
        basis, w = pins.dec(blk["basis"]), pins.dec(blk["weights"])
        side = dim * h
        X = [[Fraction(0)] * side for _ in range(side)]
        p = 0
        for s in range(dim):
            for r in range(s + 1):
                for k in range(npts):
                    v = basis[k * h:(k + 1) * h]
                    coef = x[blk["column_start"] + p] * w[p] * (Fraction(1) if r == s else Fraction(1, 2))
                    for a in range(h):
                        for b in range(h):
                            X[r * h + a][s * h + b] += coef * v[a] * v[b]
                            if r != s:
                                X[s * h + b][r * h + a] += coef * v[a] * v[b]
                    p += 1
        row = blk["row_start"]
        for j in range(side):
            for i in range(j + 1):
                if i == j:
                    out[row] += X[i][j]
                else:
                    sq[row] += X[i][j]
                row += 1
    return out, sq


def test_reference_fixture_and_known_answers(tmp_path):
    write(tmp_path, FIXTURE)
    for nl in (1, 4):
        data = sdpb_json.read_sdpb_sampled(str(tmp_path), nl)
        assert pins.dec(np.array([data["objective_constant"]]))[0] == 3
        assert data["num_equalities"] == 1
        assert (data["grams"][0]["row_start"], data["grams"][0]["side"]) == (1, 2)
        assert data["cones"] == [("zero", 1), ("psd", 2)]
        assert list(data["A"]["colptr"]) == [0, 1, 1, 2] and list(data["A"]["rowval"]) == [0, 0]
        # native_input.rs:184-198: A [1, 2, 3] = [4, -4, -4 sqrt 2, -12]
        ax, sq = apply_operator(data, [1, 2, 3])
        assert ax == [4, -4, 0, -12] and sq == [0, 0, -4, 0]


def test_reference_input_checks(tmp_path):
    write(tmp_path, FIXTURE)
    ctl = dict(FIXTURE["control.json"])
    del ctl["command"]
    write(tmp_path, {"control.json": ctl})  # the command is optional
    sdpb_json.read_sdpb_sampled(str(tmp_path), 1)
    write(tmp_path, {"unexpected.json": "{}"})
    with pytest.raises(sdpb_json.SampledInputError, match="unexpected"):
        sdpb_json.read_sdpb_sampled(str(tmp_path), 1)
    (tmp_path / "unexpected.json").unlink()
    data = dict(FIXTURE["block_data_0.json"])
    for bad, msg in ((dict(data, extra=0), "unknown field"),
                     (dict(data, bilinear_bases_even=[[2]]), "decimal"),  # a number, not a string
                     (dict(data, bilinear_bases_even=[["inf"]]), "decimal"),
                     (dict(data, bilinear_bases_even=[["1e-999999999999"]]), "underflow"),
                     (dict(data, bilinear_bases_even=[["1e999999999999"]]), "finite"),
                     (dict(data, B=[["1"], ["0"], ["1", "2"]]), "B width"),
                     (dict(data, bilinear_bases_even=[["2", "3"]]), "basis width"),
                     (dict(data, c=["1", "0"]), "inconsistent")):
        write(tmp_path, {"block_data_0.json": bad})
        with pytest.raises(sdpb_json.SampledInputError, match=msg):
            sdpb_json.read_sdpb_sampled(str(tmp_path), 4)


@pytest.mark.parametrize("nl", [1, 4, 8])
def test_decimal_rounding_is_exact_and_ties_to_even(nl):
    p = 64 * nl
    for text in ("0.1", "-2.5e-7", "3.14159265358979323846264338327950288419716939937510", "1e300", "7e-310",
                 ".5", "+12.", "-0"):
        limbs, kind, exponent = sdpb_json.decimal_record(text, nl)
        rec = mp.zeros(1, nl)
        rec[0]["limbs"] = np.array(limbs, np.uint64)
        rec[0]["kind"], rec[0]["exponent"] = kind, exponent
        assert pins.dec(rec)[0] == pins.rn(Fraction(text), p), text
    # halfway cases: 2^p + 1 rounds down to 2^p, 2^p + 3 up to 2^p + 4
    for k, want in ((1, 0), (3, 4)):
        limbs, kind, exponent = sdpb_json.decimal_record(str((1 << p) + k), nl)
        rec = mp.zeros(1, nl)
        rec[0]["limbs"] = np.array(limbs, np.uint64)
        rec[0]["kind"], rec[0]["exponent"] = kind, exponent
        assert pins.dec(rec)[0] == (1 << p) + want


def scalar_directory(path, prob):
    """A pmp2sdp-style directory of scalar blocks (dim 1, one basis parity)
    carrying a sampled problem's bases and couplings as the exact decimal
    expansions of their binary64 values (read back exactly at 256 bits)."""
    from decimal import Decimal

    def d(v):
        return str(Decimal(float(v)))

    t = prob["t"]
    write(path, {"control.json": {"num_blocks": len(prob["blocks"]), "command": "pmp2sdp"},
                 "objectives.json": {"constant": "0", "b": [d(v) for v in np.linspace(-1, 1, t)]}})
    for i, b in enumerate(prob["blocks"]):
        h, k = b["h"], b["k"]
        write(path, {f"block_info_{i}.json": {"dim": 1, "num_points": k},
                     f"block_data_{i}.json": {
                         "bilinear_bases_even": [[d(b["q_f64"][r, c]) for c in range(k)] for r in range(h)],
                         "bilinear_bases_odd": [],
                         "c": ["1"] * k,
                         "B": [[d(b["b_f64"][c, j]) for j in range(t)] for c in range(k)]}})


def test_scalar_directory_reads_back_into_the_arrow_kkt(tmp_path):
    nl = 4
    prob = workloads.sampled_problem(nl, nblocks=3, hmin=4, hmax=7, t=9, seed=13, lift=False)
    scalar_directory(tmp_path, prob)
    parsed = sdpb_json.read_sdpb_sampled(str(tmp_path), nl)
    arrow = sdpb_json.arrow_problem(parsed, nl, rinv=[b["rinv"] for b in prob["blocks"]],
                                    weights=[b["w"] for b in prob["blocks"]], border=prob["border"])
    for got, want in zip(arrow["blocks"], prob["blocks"]):
        assert (got["h"], got["k"]) == (want["h"], want["k"])
        for key in ("q", "b"):
            assert np.all(mp.same_bits(got[key], want[key])), key
    with pytest.raises(sdpb_json.SampledInputError, match="dim 1"):
        bad = dict(parsed, sampled=[dict(parsed["sampled"][0], dim=2)])
        sdpb_json.arrow_problem(bad, nl)
