"""The oracles pinned before they are trusted.

* mpexact (plain Python integers) against the reference's known answers:
  the cancellation case of tests/it/mpfr_gemm_syrk.rs:57-61,112-117, and
  rounding against an independent rational rounding and mpmath.
* oracle.c (the C restatement) against mpexact on the reference's LCG test
  operands (rns_blas_tests.rs:137-145) and against the committed fixture.
* A planted mistake in the C oracle's caller (a transposed operand) must be
  caught by the same comparison.
"""
import hashlib
import json
import os
import random
from fractions import Fraction

import mpmath
import numpy as np
import pytest

import mpexact as M
from helpers import from_mp, lcg_operands, oracle_entries, oracle_gemm, to_mp
from sdpx_gpu import mp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rns_gemm.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("nl", [2, 4, 8, 12, 16])
def test_cancellation_known_answer(nl):
    # -1 + (1+d)(1-d) = -d^2, d = 2^-(32N+1): lost by separate roundings,
    # kept by exact accumulation (mpfr_gemm_syrk.rs:57-61, 112-117).
    d = M.round_fraction(Fraction(1, 2 ** (32 * nl + 1)), nl)
    one = M.from_int(1, nl)
    v = M.exact_dot([(M.neg(one), one), (M.add(one, d), M.sub(one, d))], nl)
    assert v.to_fraction() == -Fraction(1, 2 ** (64 * nl + 2))
    assert v.kind == -M.REGULAR_KIND and v.exponent == -(64 * nl + 1)


def test_rounding_matches_rational_and_mpmath():
    rnd = random.Random(5)
    for _ in range(400):
        nl = rnd.choice([1, 2, 4])
        p = 64 * nl
        bits = rnd.randint(1, 3 * p)
        mag = rnd.getrandbits(bits) | (1 << (bits - 1))
        if rnd.random() < 0.3 and bits > p + 1:
            # force an exact tie: round bit 1, sticky 0
            shift = bits - p
            mag = (mag >> shift << shift) | (1 << (shift - 1))
        scale = rnd.randint(-500, 500)
        neg = rnd.random() < 0.5
        a = M.from_scaled_integer(neg, mag, scale, nl)
        b = M.round_fraction(Fraction(-mag if neg else mag) * Fraction(2) ** scale, nl)
        assert a.key() == b.key()
        with mpmath.workprec(p):
            ref = mpmath.ldexp(mpmath.mpf(-mag if neg else mag), scale)
        sign, man, ex, _ = ref._mpf_
        assert Fraction((-1) ** sign * man) * Fraction(2) ** ex == a.to_fraction()


def test_known_values():
    # 1/3 at 64 bits: 0xAAAA...AB * 2^-65 (nearest-even rounds up).
    v = M.round_fraction(Fraction(1, 3), 1)
    assert v.limbs == (0xAAAAAAAAAAAAAAAB,) and v.exponent == -1
    # 1.0, 0.5, 3.0 layouts (dyadic.rs:10-11 pins these limb by limb).
    assert M.from_int(1, 2).limbs == (0, 1 << 63) and M.from_int(1, 2).exponent == 1
    assert M.from_int(3, 1).limbs == (0xC000000000000000,) and M.from_int(3, 1).exponent == 2


@pytest.mark.parametrize("nl", [1, 2, 8])
def test_from_float64_is_exact(nl):
    # binary64 -> records (subnormals, signed zeros, extremes, random
    # magnitudes over the whole range): the record's exact value is the
    # double's, and the layout is the general integer conversion's
    rng = np.random.default_rng(nl)
    x = np.concatenate([[0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                         1.0, -0.5, 3.0, 1 / 3],
                        rng.standard_normal(300) * np.exp(rng.uniform(-700, 700, 300))])
    rec = mp.from_float64(x, nl)
    for v, r in zip(x, to_mp(rec)):
        assert r.to_fraction() == Fraction(float(v))
    mant, ex = np.frexp(np.abs(x))
    gen = mp.from_integers((mant * 2.0**53).astype(np.uint64)[:, None], x < 0, ex.astype(np.int64) - 53, nl)
    assert rec.tobytes() == gen.tobytes()


def test_mpexact_gemm_brute_force_tiny():
    # exact rational brute force on 2x3x2 with mixed exponents
    rnd = random.Random(11)
    nl = 2
    vals = [M.round_fraction(Fraction(rnd.randint(-10**30, 10**30), 2 ** rnd.randint(0, 200)), nl)
            for _ in range(12)]
    a, b = vals[:6], vals[6:]
    out = M.gemm_exact("N", "N", 2, 2, 3, a, 2, b, 3)
    for i in range(2):
        for j in range(2):
            s = sum(a[i + p * 2].to_fraction() * b[p + j * 3].to_fraction() for p in range(3))
            assert out[i + j * 2].key() == M.round_fraction(s, nl).key()


def test_golden_fixture_matches_mpexact_inputs_and_c_oracle():
    g = golden()
    seen = set()
    for case in g["cases"]:
        key = (case["limbs"], case["m"], case["n"], case["k"], case["seed"])
        if key in seen:
            continue
        seen.add(key)
        nl, m, n, k = case["limbs"], case["m"], case["n"], case["k"]
        rng = M.Lcg(case["seed"])
        for c2 in [c for c in g["cases"] if (c["limbs"], c["m"], c["n"], c["k"], c["seed"]) == key]:
            ta, tb = c2["ta"], c2["tb"]
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            a = from_mp(M.lcg_random(rng, ar * ac, c2["spread"], nl), nl)
            b = from_mp(M.lcg_random(rng, br * bc, c2["spread"], nl), nl)
            assert hashlib.sha256(a.tobytes()).hexdigest() == c2["a_sha256"]
            assert hashlib.sha256(b.tobytes()).hexdigest() == c2["b_sha256"]
            out = oracle_gemm(ta, tb, m, n, k, a, ar, b, br)
            assert hashlib.sha256(out.tobytes()).hexdigest() == c2["out_sha256"]
            if m == n:
                packed = oracle_gemm(ta, tb, m, n, k, a, ar, b, br, upper=True)
                assert hashlib.sha256(packed.tobytes()).hexdigest() == c2["packed_sha256"]


def test_cancellation_fixture():
    for kat in golden()["cancellation"]:
        nl = kat["limbs"]
        assert kat["kind"] == -3 and kat["exponent"] == -(64 * nl + 1)
        assert [int(x, 16) for x in kat["limbs_hex"]] == [0] * (nl - 1) + [1 << 63]


@pytest.mark.parametrize("nl,m,n,k,spread", [(4, 13, 11, 37, 60), (2, 6, 9, 70, 5), (8, 8, 8, 17, 300)])
def test_c_oracle_matches_mpexact(nl, m, n, k, spread):
    for ta, tb in (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")):
        a, lda, b, ldb = lcg_operands(m, n, k, spread, 99 + m, nl, ta, tb)
        ref = M.gemm_exact(ta, tb, m, n, k, to_mp(a), lda, to_mp(b), ldb)
        got = oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb)
        assert np.all(mp.same_bits(got, from_mp(ref, nl)))
        # spot entries agree with the full product
        rows = np.array([0, m - 1, m // 2]); cols = np.array([n - 1, 0, n // 2])
        ent = oracle_entries(ta, tb, k, a, lda, b, ldb, rows, cols)
        assert np.all(mp.same_bits(ent, got[rows + cols * m]))


def test_planted_transpose_is_caught():
    nl, m, n, k = 4, 7, 7, 19
    a, lda, b, ldb = lcg_operands(m, n, k, 30, 3, nl, "N", "N")
    ref = from_mp(M.gemm_exact("N", "N", m, n, k, to_mp(a), lda, to_mp(b), ldb), nl)
    wrong = oracle_gemm("T", "N", m, n, k, a, lda, b, ldb) if m == k else None
    # a transposed A (7x19 read as 19x7 is out of shape), so plant by
    # swapping operand roles instead: B^T A^T != A B in general
    wrong = oracle_gemm("T", "T", n, m, k, b, ldb, a, lda)
    wrong_t = wrong.reshape(n, m).T.reshape(-1)  # (B^T A^T)^T laid out as m x n
    assert np.all(mp.same_bits(wrong_t, ref))  # identity holds exactly ...
    planted = oracle_gemm("N", "N", m, n, k, a, lda, b[::-1].copy(), ldb)
    assert not np.all(mp.same_bits(planted, ref))  # ... a wrong operand does not


def test_decline_rules():
    nl = 4
    a = mp.random_full(12, nl, 10, np.random.default_rng(1))
    b = mp.random_full(12, nl, 10, np.random.default_rng(2))
    assert oracle_gemm("N", "N", 3, 3, 4, a, 3, b, 4) is not None
    bad = a.copy(); bad["kind"][5] = M.INF_KIND
    assert oracle_gemm("N", "N", 3, 3, 4, bad, 3, b, 4) is None
    bad = a.copy(); bad["kind"][5] = M.NAN_KIND
    assert oracle_gemm("N", "N", 3, 3, 4, bad, 3, b, 4) is None
    wide = a.copy(); wide["kind"][:] = 3; wide["exponent"][0] = 5000; wide["exponent"][1] = -100
    assert oracle_gemm("N", "N", 3, 3, 4, wide, 3, b, 4) is None
    ok = a.copy(); ok["kind"][:] = 3; ok["exponent"][:] = 0; ok["exponent"][0] = 4096
    assert oracle_gemm("N", "N", 3, 3, 4, ok, 3, b, 4) is not None


# ---------------------------------------------------------------------------
# Exact-dot Cholesky / triangular solves (mpfr.rs:726-872, traits.rs:85-113)
# ---------------------------------------------------------------------------

def _mp_of(m, e, n, neg=False):
    return M.Mp(n, [(m >> (64 * i)) & ((1 << 64) - 1) for i in range(n)], -3 if neg else 3, e)


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_sqrt_and_division_correctly_rounded(n):
    # closed form: mpmath's correctly rounded sqrt / division at 64n bits
    from mpmath import libmp
    rng = random.Random(n)
    p = 64 * n
    specials = [1 << (p - 1), (1 << p) - 1, (1 << (p - 1)) + 1, 3 << (p - 2)]
    for t in range(200):
        m = specials[t] if t < len(specials) else rng.getrandbits(p) | (1 << (p - 1))
        e = rng.randint(-400, 400)
        x = _mp_of(m, e, n)
        want = libmp.mpf_sqrt(libmp.from_man_exp(m, e - p), p, "n")
        assert M.sqrt(x).to_fraction() == Fraction(*libmp.to_rational(want))
        m2 = rng.getrandbits(p) | (1 << (p - 1))
        e2 = rng.randint(-300, 300)
        y = _mp_of(m2, e2, n, neg=bool(t & 1))
        want = libmp.mpf_div(libmp.from_man_exp(m, e - p), libmp.from_man_exp(-m2 if t & 1 else m2, e2 - p), p, "n")
        assert M.div(x, y).to_fraction() == Fraction(*libmp.to_rational(want))


@pytest.mark.parametrize("n", [2, 4, 8])
def test_c_oracle_sqrt_div_match_python(n):
    from helpers import oracle
    rng = np.random.default_rng(n)
    a = mp.random_full(60, n, 300, rng, zero_fraction=0.0)
    a["kind"] = np.abs(a["kind"])  # positive radicands
    b = mp.random_full(60, n, 300, rng, zero_fraction=0.0)
    L = oracle()
    out = mp.zeros(1, n)
    for i in range(60):
        L.oracle_sqrt(n, a[i:i + 1].ctypes.data, out.ctypes.data)
        assert to_mp(out)[0].key() == M.sqrt(to_mp(a[i:i + 1])[0]).key()
        L.oracle_div(n, a[i:i + 1].ctypes.data, b[i:i + 1].ctypes.data, out.ctypes.data)
        assert to_mp(out)[0].key() == M.div(to_mp(a[i:i + 1])[0], to_mp(b[i:i + 1])[0]).key()


def test_potrf_potrs_reference_known_answer():
    # tests/it/mpfr_dense.rs:29-38: xpotrf('L', 2, [4, 2, 987, 3]) keeps the
    # upper entry, and xpotrs solves A x = (6, 5) to x = (1, 1)
    from helpers import oracle_potrf, oracle_potrs
    for nl in (2, 4):
        a = mp.from_float64(np.array([4.0, 2.0, 987.0, 3.0]), nl)
        info, ch = oracle_potrf(a, 2, 2, "L")
        assert info == 0
        assert mp.to_float64(ch)[2] == 987.0
        py = list(to_mp(a))
        assert M.potrf("L", 2, py, 2) == 0
        assert [v.key() for v in py] == [v.key() for v in to_mp(ch)]
        rhs = oracle_potrs(ch, 2, 2, mp.from_float64(np.array([6.0, 5.0]), nl), 1, 2, "L")
        x = [v.to_fraction() for v in to_mp(rhs)]
        eps = Fraction(1, 2 ** (64 * nl - 1))
        assert all(abs(v - 1) <= 10000 * eps for v in x)


def _spd(n, rng, nl, spread=20):
    g = rng.standard_normal((n, n)) * 2.0 ** rng.integers(-spread, spread, (n, n))
    s = g @ g.T + n * np.eye(n)
    return mp.from_float64(s.T.reshape(-1), nl), s


@pytest.mark.parametrize("nl,n,uplo", [(2, 9, "L"), (2, 7, "U"), (4, 6, "L")])
def test_c_oracle_potrf_trsm_potrs_match_python(nl, n, uplo):
    from helpers import oracle_potrf, oracle_potrs, oracle_trsm_lower
    rng = np.random.default_rng(n)
    a, _ = _spd(n, rng, nl)
    info, ch = oracle_potrf(a, n, n, uplo)
    py = list(to_mp(a))
    assert info == 0 and M.potrf(uplo, n, py, n) == 0
    assert [v.key() for v in py] == [v.key() for v in to_mp(ch)]
    nrhs = 3
    b = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    x = oracle_potrs(ch, n, n, b, nrhs, n, uplo)
    pyb = list(to_mp(b))
    M.potrs(uplo, n, nrhs, py, n, pyb, n)
    assert [v.key() for v in pyb] == [v.key() for v in to_mp(x)]
    if uplo == "L":
        for tr in (False, True):
            t = oracle_trsm_lower(ch, n, b, tr)
            pyt = list(to_mp(b))
            M.trsm_lower(n, py, pyt, tr)
            assert [v.key() for v in pyt] == [v.key() for v in to_mp(t)]


def test_potrf_reports_failing_pivot():
    # an indefinite matrix: info = failing column + 1, the diagonal keeps the
    # (non-positive) pivot before the square root (mpfr.rs:745-750)
    from helpers import oracle_potrf
    s = np.array([[4.0, 2.0, 0.0], [2.0, 1.0, 1.0], [0.0, 1.0, 5.0]])
    a = mp.from_float64(s.T.reshape(-1), 2)
    info, ch = oracle_potrf(a, 3, 3, "L")
    assert info == 2
    assert mp.to_float64(ch)[4] == 0.0
    py = list(to_mp(a))
    assert M.potrf("L", 3, py, 3) == 2


@pytest.mark.parametrize("nl,m,n,k,spread", [(4, 7, 5, 45, 768), (4, 20, 20, 20, 768), (2, 3, 4, 1, 512),
                                             (4, 9, 6, 300, 256), (8, 35, 37, 45, 1024), (4, 64, 64, 70, 4)])
def test_cpu_residue_product_matches_exact_dots(nl, m, n, k, spread):
    """The CPU residue product (primes of 18-26 bits in binary64, CRT, one
    rounding: the reference's CPU algorithm, used as the CPU baseline's
    product) gives the exact dots' bits on the reference's test shapes and
    spreads (rns_blas_tests.rs:137-145), every transpose, full and packed
    upper; a planted operand change is seen."""
    from helpers import oracle_gemm_rns

    for ta in "NT":
        for tb in "NT":
            a, lda, b, ldb = lcg_operands(m, n, k, spread, 7 + m + k, nl, ta, tb)
            want = oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb)
            got = oracle_gemm_rns(ta, tb, m, n, k, a, lda, b, ldb)
            assert got is not None and np.all(mp.same_bits(got, want)), (ta, tb)
            if m == n:
                want = oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb, upper=True)
                got = oracle_gemm_rns(ta, tb, m, n, k, a, lda, b, ldb, upper=True)
                assert np.all(mp.same_bits(got, want)), (ta, tb, "upper")
    a2 = a.copy()
    top = np.argmax(a["exponent"] * (np.abs(a["kind"]) == 3))  # the largest entry's high bits
    a2[top]["limbs"][nl - 1] ^= np.uint64(1 << 40)
    assert not np.all(mp.same_bits(oracle_gemm_rns(ta, tb, m, n, k, a2, lda, b, ldb),
                                   oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb)))
