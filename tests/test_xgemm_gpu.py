"""sdpx_gpu_xgemm / sdpx_gpu_xsyrk (the MPFR xgemm and xsyrk with alpha and
beta, blas/mpfr.rs:245-348, 540-626) on the GPU, bit for bit against the C
oracle's exact product followed by the oracle's axpby (mpfr.rs:61-80)."""
from fractions import Fraction

import numpy as np
import pytest

import mpexact
import pins
from helpers import from_mp, lcg_operands, oracle_gemm, to_mp
from sdpx_gpu import mp
from test_exact_forms import axpby_case

pytestmark = pytest.mark.gpu

SCALARS = {"zero": Fraction(0), "one": Fraction(1), "minus_one": Fraction(-1), "other": Fraction(-5, 7)}


def scalar(name, nl):
    return pins.enc([pins.rn(SCALARS[name], 64 * nl)], nl)


def test_axpby_known_answer_gpu(gpu):
    nl = 2
    alpha, x, beta, y, want, _ = axpby_case(64 * nl)
    c = pins.enc([y], nl)
    assert gpu.xgemm("N", "N", 1, 1, 1, pins.enc([alpha], nl), pins.enc([x], nl), 1, pins.enc([1], nl), 1,
                     pins.enc([beta], nl), c, 1)
    assert pins.dec(c) == [want]


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [("one", "zero"), ("other", "other"), ("minus_one", "one"),
                                        ("zero", "minus_one"), ("one", "other")])
def test_xgemm_vs_oracle(gpu, nl, ta, tb, alpha, beta):
    m, n, k, ldc = 37, 29, 45, 41
    print("xgemm", nl, ta, tb, alpha, beta, flush=True)
    a, lda, b, ldb = lcg_operands(m, n, k, 30, 5 + nl, nl, ta, tb)
    c0 = from_mp(mpexact.lcg_random(mpexact.Lcg(77), ldc * n, 30, nl), nl)
    al, be = scalar(alpha, nl), scalar(beta, nl)
    c = c0.copy()
    assert gpu.xgemm(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc)
    prod = to_mp(oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb))
    alv, bev, cv = to_mp(al)[0], to_mp(be)[0], to_mp(c0)
    got = to_mp(c)
    for j in range(n):
        for i in range(ldc):
            e = i + j * ldc
            if i >= m:
                assert c[e].tobytes() == c0[e].tobytes()
                continue
            want = mpexact.axpby(alv, prod[i + j * m], bev, cv[e])
            assert got[e].key() == want.key(), (i, j)


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("uplo,trans", [("U", "N"), ("U", "T"), ("L", "N"), ("L", "T")])
@pytest.mark.parametrize("packed", [False, True])
def test_xsyrk_vs_oracle(gpu, nl, uplo, trans, packed):
    if packed and uplo == "L":
        pytest.skip("only the upper triangle packs")
    n, k = 33, 50
    ldc = n if packed else n + 3
    print("xsyrk", nl, uplo, trans, packed, flush=True)
    a, lda, _, _ = lcg_operands(n, n, k, 30, 11 + nl, nl, trans, "N")
    clen = n * (n + 1) // 2 if packed else ldc * n
    c0 = from_mp(mpexact.lcg_random(mpexact.Lcg(78), clen, 30, nl), nl)
    al, be = scalar("other", nl), scalar("other", nl)
    c = c0.copy()
    assert gpu.xsyrk(uplo, trans, n, k, al, a, lda, be, c, ldc)
    prod = to_mp(oracle_gemm(trans, "T" if trans == "N" else "N", n, n, k, a, lda, a, lda))
    alv, bev, cv, got = to_mp(al)[0], to_mp(be)[0], to_mp(c0), to_mp(c)
    for j in range(n):
        for i in range(ldc if not packed else j + 1):
            e = j * (j + 1) // 2 + i if packed else i + j * ldc
            inside = i < n and ((i <= j) if uplo == "U" else (i >= j))
            if not inside:
                assert c[e].tobytes() == c0[e].tobytes(), (i, j)
                continue
            want = mpexact.axpby(alv, prod[i + j * n], bev, cv[e])
            assert got[e].key() == want.key(), (i, j)


def test_xgemm_alpha_zero_skips_the_product(gpu):
    """alpha = 0 reads neither a nor b (they may even be non-finite), as the
    reference's gemm, and c = axpby(0, 0, beta, c)."""
    nl, m, n, k = 2, 4, 3, 5
    nan = mp.zeros(m * k, nl)
    nan["kind"] = 0
    c0 = pins.enc([Fraction(2 * i + 1, 8) for i in range(m * n)], nl)
    c = c0.copy()
    assert gpu.xgemm("N", "N", m, n, k, scalar("zero", nl), nan, m, nan, k, scalar("minus_one", nl), c, m)
    assert pins.dec(c) == [-v for v in pins.dec(c0)]
