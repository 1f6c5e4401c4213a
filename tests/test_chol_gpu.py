"""Exact-dot Cholesky, solve and triangular solve on the GPU (xpotrf,
xpotrs: blas/mpfr.rs:727-872; xtrsm_lower: blas/traits.rs:85-113), bitwise
against the C oracle, itself pinned in test_oracle.py (known answer of
tests/it/mpfr_dense.rs:29-38, the Python restatement, the failing-pivot
rule)."""
import numpy as np
import pytest

import sdpx_gpu
from helpers import oracle_potrf, oracle_potrs, oracle_trsm_lower
from sdpx_gpu import mp

pytestmark = pytest.mark.gpu


def spd(n, rng, nl, spread=20):
    """G G^T + n I with entries of G scaled by 2^U{-spread..spread}."""
    g = rng.standard_normal((n, n)) * 2.0 ** rng.integers(-spread, spread, (n, n))
    s = g @ g.T + n * np.eye(n)
    return mp.from_float64(s.T.reshape(-1), nl)


def padded(a, n, lda, rng, nl):
    """Column-major copy with leading dimension lda; padding rows hold
    random values that must survive untouched."""
    out = mp.from_float64(rng.standard_normal(lda * n), nl)
    for c in range(n):
        out[c * lda:c * lda + n] = a[c * n:(c + 1) * n]
    return out


def test_reference_known_answer(gpu):
    # mpfr_dense.rs:29-38: xpotrf('L', 2, [4, 2, 987, 3]) keeps the upper
    # entry; xpotrs solves A x = (6, 5) to x = (1, 1)
    for nl in (2, 4, 8):
        a = mp.from_float64(np.array([4.0, 2.0, 987.0, 3.0]), nl)
        info_o, ch_o = oracle_potrf(a, 2, 2, "L")
        ch = a.copy()
        assert sdpx_gpu.xpotrf("L", 2, ch, 2) == 0 == info_o
        assert mp.to_float64(ch)[2] == 987.0
        assert np.all(mp.same_bits(ch, ch_o))
        b = mp.from_float64(np.array([6.0, 5.0]), nl)
        x = b.copy()
        assert sdpx_gpu.xpotrs("L", 2, 1, ch, 2, x, 2) == 0
        assert np.all(mp.same_bits(x, oracle_potrs(ch_o, 2, 2, b, 1, 2, "L")))
        assert np.allclose(mp.to_float64(x), [1.0, 1.0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("nl,n,uplo,pad", [(2, 40, "L", 0), (2, 37, "U", 3), (4, 96, "L", 5),
                                           (4, 70, "U", 0), (8, 130, "L", 0), (8, 64, "U", 2)])
def test_potrf_potrs_match_oracle(gpu, nl, n, uplo, pad):
    print("potrf", nl, n, uplo, pad, flush=True)
    rng = np.random.default_rng(1000 * nl + n)
    lda = n + pad
    a = padded(spd(n, rng, nl), n, lda, rng, nl)
    info_o, ch_o = oracle_potrf(a, n, lda, uplo)
    ch = a.copy()
    info = sdpx_gpu.xpotrf(uplo, n, ch, lda)
    assert info == info_o == 0
    assert np.all(mp.same_bits(ch, ch_o))
    nrhs, ldb = 5, n + 1
    b = mp.from_float64(rng.standard_normal(ldb * nrhs) * 2.0 ** rng.integers(-30, 30, ldb * nrhs), nl)
    x = b.copy()
    assert sdpx_gpu.xpotrs(uplo, n, nrhs, ch, lda, x, ldb) == 0
    assert np.all(mp.same_bits(x, oracle_potrs(ch_o, n, lda, b, nrhs, ldb, uplo)))


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_failing_pivot_matches_oracle(gpu, uplo):
    # an indefinite matrix: info names the first failing column, its pivot
    # d is stored unrooted and the columns after it are untouched
    nl, n = 4, 24
    rng = np.random.default_rng(7)
    s = rng.standard_normal((n, n))
    s = s @ s.T + n * np.eye(n)
    s[15, 15] = -5.0 * n
    a = mp.from_float64(s.T.reshape(-1), nl)
    info_o, ch_o = oracle_potrf(a, n, n, uplo)
    ch = a.copy()
    info = sdpx_gpu.xpotrf(uplo, n, ch, n)
    assert info == info_o and 1 <= info <= 16
    assert np.all(mp.same_bits(ch, ch_o))


@pytest.mark.parametrize("nl,n,nrhs", [(2, 50, 3), (4, 120, 7), (8, 90, 2)])
def test_trsm_lower_matches_oracle(gpu, nl, n, nrhs):
    print("trsm_lower", nl, n, nrhs, flush=True)
    rng = np.random.default_rng(n + nl)
    # bounded growth: off-diagonal entries 2^U{-12..-3} N(0,1), diagonal in
    # [1, 2); the solved entries then span a few hundred bits at most
    lo = np.tril(rng.standard_normal((n, n)) * 2.0 ** rng.integers(-12, -3, (n, n)), -1)
    lo[np.diag_indices(n)] = 1.0 + rng.random(n)
    a = mp.from_float64(lo.T.reshape(-1), nl)
    b = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    for tr in (False, True):
        x = b.copy()
        sdpx_gpu.xtrsm_lower(n, a, x, tr)
        assert np.all(mp.same_bits(x, oracle_trsm_lower(a, n, b, tr))), tr


def test_argument_codes_and_declines(gpu):
    nl, n = 2, 4
    a = mp.from_float64(np.eye(n).reshape(-1), nl)
    assert sdpx_gpu.xpotrf("X", n, a.copy(), n) == -1
    assert sdpx_gpu.xpotrf("L", n, a.copy(), n - 1) == -4
    b = mp.from_float64(np.ones(n), nl)
    assert sdpx_gpu.xpotrs("L", n, 1, a, n, b.copy(), n - 1) == -7
    # a non-finite entry: declined, nothing written
    bad = a.copy()
    bad[1]["kind"] = 0
    before = bad.copy()
    with pytest.raises(sdpx_gpu.DeviceError):
        sdpx_gpu.xpotrf("L", n, bad, n)
    assert np.all(mp.same_bits(bad, before))


def test_trsm_lower_growth_matches_or_declines(gpu):
    # entries up to 2^20 make the solution grow by ~2^20 a row: the exact
    # dots then span more exponents than the GPU window holds, and the call
    # declines (code 2) with b untouched, or it matches the oracle bitwise
    nl, n = 4, 120
    rng = np.random.default_rng(3)
    lo = np.tril(rng.standard_normal((n, n)) * 2.0 ** rng.integers(-20, 20, (n, n)))
    lo[np.diag_indices(n)] = 1.0 + rng.random(n)
    a = mp.from_float64(lo.T.reshape(-1), nl)
    b = mp.from_float64(rng.standard_normal(n), nl)
    x = b.copy()
    try:
        sdpx_gpu.xtrsm_lower(n, a, x, False)
    except sdpx_gpu.DeviceError as e:
        assert "error 2" in str(e)
        assert np.all(mp.same_bits(x, b))
        return
    assert np.all(mp.same_bits(x, oracle_trsm_lower(a, n, b, False)))
