"""Writes tests/golden/rns_gemm.json: expected outputs of the exact product
for the reference's own residue-BLAS test cases.

Inputs are the reference's LCG operands (rns_blas_tests.rs:4-28, 62-70;
shapes, precisions, spreads and seeds of
residue_blas_gemm_matches_exact_dots_bitwise, rns_blas_tests.rs:137-145),
regenerated exactly by oracle/mpexact.py.  Expected outputs come from
mpexact's big-integer exact dots rounded once (the reference's `reference`
function, rns_blas_tests.rs:30-60, computes the same exact dots).  The
fixture stores SHA-256 digests of the output bytes plus the first entries
in full, so it stays small.

Run from the repo root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "sdpx.rs_amd"),
                os.path.join(ROOT, "tests")]
import mpexact  # noqa: E402
from helpers import from_mp  # noqa: E402

# (limbs, m, n, k, spread, seed): rns_blas_tests.rs:139-144
CASES = [(12, 7, 5, 45, 90, 1), (12, 20, 20, 20, 3, 2), (8, 3, 4, 1, 40, 3),
         (4, 9, 6, 300, 60, 4), (2, 5, 5, 33, 10, 5), (16, 35, 37, 45, 90, 7)]
KEEP = 4


def digest(rec):
    return hashlib.sha256(rec.tobytes()).hexdigest()


def main():
    out = []
    for nl, m, n, k, spread, seed in CASES:
        rng = mpexact.Lcg(seed)
        for ta, tb in (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")):
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            a = mpexact.lcg_random(rng, ar * ac, spread, nl)
            b = mpexact.lcg_random(rng, br * bc, spread, nl)
            full = mpexact.gemm_exact(ta, tb, m, n, k, a, ar, b, br)
            rec = from_mp(full, nl)
            entry = {"limbs": nl, "m": m, "n": n, "k": k, "spread": spread, "seed": seed,
                     "ta": ta, "tb": tb, "lda": ar, "ldb": br,
                     "a_sha256": digest(from_mp(a, nl)), "b_sha256": digest(from_mp(b, nl)),
                     "out_sha256": digest(rec),
                     "first": [[str(v.kind), str(v.exponent), [str(x) for x in v.limbs]]
                               for v in full[:KEEP]]}
            if m == n:
                packed = mpexact.gemm_exact(ta, tb, m, n, k, a, ar, b, br, upper_only=True)
                entry["packed_sha256"] = digest(from_mp(packed, nl))
            out.append(entry)
            print(nl, m, n, k, ta, tb, flush=True)
    # The cancellation known answer of tests/it/mpfr_gemm_syrk.rs:57-61, 112-117:
    # -1*1 + (1+d)(1-d) + 0*0 with d = 2^-(32N+1) keeps -d^2.
    kat = []
    for nl in (4, 8, 12, 16):
        d = mpexact.round_fraction(mpexact.Fraction(1, 2 ** (32 * nl + 1)), nl)
        one = mpexact.from_int(1, nl)
        v = mpexact.exact_dot([(mpexact.neg(one), one), (mpexact.add(one, d), mpexact.sub(one, d))], nl)
        kat.append({"limbs": nl, "kind": v.kind, "exponent": v.exponent,
                    "limbs_hex": [hex(x) for x in v.limbs]})
    with open(os.path.join(ROOT, "tests", "golden", "rns_gemm.json"), "w") as f:
        json.dump({"cases": out, "cancellation": kat}, f, indent=1)


if __name__ == "__main__":
    main()
