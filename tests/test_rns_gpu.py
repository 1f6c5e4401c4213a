"""Exact RNS products on the GPU, bit-exact against the oracles.

Covers the reference's residue-BLAS test cases (rns_blas_tests.rs:137-145,
via the committed digests), all transposes, packed upper outputs, the
same-operand Gram (SYRK) form, long inner dimensions (k = 70000 as in
rns_blas_tests.rs:226-229), zero operands, decline rules
(rns_blas_tests.rs:476-495 forces Inf), and a spot check at the full
4096^3 workload size.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import mpexact as M
from helpers import from_mp, oracle_entries, oracle_gemm
from sdpx_gpu import mp

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rns_gemm.json")
TRANS = (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T"))


def sha(rec):
    return hashlib.sha256(rec.tobytes()).hexdigest()


def test_golden_cases(gpu):
    with open(GOLDEN) as f:
        g = json.load(f)
    rngs = {}
    for c in g["cases"]:
        key = (c["limbs"], c["m"], c["n"], c["k"], c["seed"])
        rng = rngs.setdefault(key, M.Lcg(c["seed"]))
        nl, m, n, k = c["limbs"], c["m"], c["n"], c["k"]
        print("golden", nl, m, n, k, c["ta"], c["tb"], flush=True)
        a = from_mp(M.lcg_random(rng, c["lda"] * (k if c["ta"] == "N" else m), c["spread"], nl), nl)
        b = from_mp(M.lcg_random(rng, c["ldb"] * (n if c["tb"] == "N" else k), c["spread"], nl), nl)
        assert sha(a) == c["a_sha256"] and sha(b) == c["b_sha256"]
        ok, out = gpu.rns_gemm(c["ta"], c["tb"], m, n, k, a, c["lda"], b, c["ldb"])
        assert ok
        assert sha(out) == c["out_sha256"], (nl, m, n, k, c["ta"], c["tb"])
        if m == n:
            ok, packed = gpu.rns_gemm(c["ta"], c["tb"], m, n, k, a, c["lda"], b, c["ldb"],
                                      upper_only=True)
            assert ok and sha(packed) == c["packed_sha256"]


@pytest.mark.parametrize("nl,m,n,k,spread", [
    (4, 300, 257, 1000, 60), (8, 129, 130, 131, 200), (2, 64, 200, 513, 8),
    (12, 40, 33, 77, 90), (16, 35, 37, 45, 90), (1, 128, 128, 128, 3),
    # operand windows just inside / outside the 768-bit matrix-core encode
    (8, 64, 65, 300, 120), (8, 64, 65, 300, 130)])
def test_random_vs_oracle(gpu, nl, m, n, k, spread):
    rng = np.random.default_rng(nl * 1000 + m)
    for ta, tb in TRANS:
        print("random", nl, m, n, k, ta, tb, flush=True)
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb = ar + 3, br + 1  # padded leading dimensions
        a = mp.random_full(lda * ac, nl, spread, rng)
        b = mp.random_full(ldb * bc, nl, spread, rng)
        ref = oracle_gemm(ta, tb, m, n, k, a, lda, b, ldb)
        ok, out = gpu.rns_gemm(ta, tb, m, n, k, a, lda, b, ldb)
        assert ok
        bad = ~mp.same_bits(out, ref)
        assert not bad.any(), f"{bad.sum()} mismatches, first at {np.argmax(bad)}"


def test_upper_unpacked_leaves_lower(gpu):
    rng = np.random.default_rng(7)
    nl, m, n, k = 4, 70, 90, 50
    a = mp.random_full(m * k, nl, 20, rng)
    b = mp.random_full(k * n, nl, 20, rng)
    ref = oracle_gemm("N", "N", m, n, k, a, m, b, k)
    sentinel = mp.random_full(m * n, nl, 5, rng)
    ok, out = gpu.rns_gemm("N", "N", m, n, k, a, m, b, k, upper_only=True, out=sentinel.copy())
    assert ok
    i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    upper = (i <= j).T.reshape(-1)  # column-major index i + j*m
    idx = np.arange(m * n)
    up = upper[(idx % m) * n + idx // m] if False else ((idx % m) <= (idx // m))
    assert np.all(mp.same_bits(out[up], ref[up]))
    assert np.all(mp.same_bits(out[~up], sentinel[~up]))


@pytest.mark.parametrize("nl,n,k", [(4, 333, 1200), (8, 150, 400)])
def test_gram_same_operand_packed(gpu, nl, n, k):
    rng = np.random.default_rng(n)
    a = mp.random_full(k * n, nl, 30, rng)
    ref = oracle_gemm("T", "N", n, n, k, a, k, a, k, upper=True)
    ok, out = gpu.rns_gemm("T", "N", n, n, k, a, k, a, k, upper_only=True)
    assert ok and np.all(mp.same_bits(out, ref))
    # A A^T form
    ref2 = oracle_gemm("N", "T", n, n, k, a.reshape(-1)[: n * k], n, a.reshape(-1)[: n * k], n,
                       upper=True)
    ok, out2 = gpu.rns_gemm("N", "T", n, n, k, a[: n * k], n, a[: n * k], n, upper_only=True)
    assert ok and np.all(mp.same_bits(out2, ref2))


def test_long_inner_dimension(gpu):
    # k = 70000 crosses two exact INT32 accumulation chunks
    rng = np.random.default_rng(3)
    nl, m, n, k = 4, 5, 6, 70000
    a = mp.random_full(m * k, nl, 4, rng)
    b = mp.random_full(k * n, nl, 4, rng)
    ref = oracle_gemm("N", "N", m, n, k, a, m, b, k)
    ok, out = gpu.rns_gemm("N", "N", m, n, k, a, m, b, k)
    assert ok and np.all(mp.same_bits(out, ref))


def test_cancellation_and_ties(gpu):
    for nl in (2, 4, 8):
        d = M.round_fraction(M.Fraction(1, 2 ** (32 * nl + 1)), nl)
        one = M.from_int(1, nl)
        a = from_mp([M.neg(one), M.add(one, d), M.zero(nl)], nl)  # 1 x 3
        b = from_mp([one, M.sub(one, d), M.zero(nl)], nl)          # 3 x 1
        ok, out = gpu.rns_gemm("N", "N", 1, 1, 3, a, 1, b, 3)
        assert ok
        assert out["kind"][0] == -3 and out["exponent"][0] == -(64 * nl + 1)
        # an exact tie: (2^P + 1) * 1 rounds to even 2^P; (2^P + 3) -> 2^P + 4
        p = 64 * nl
        big = M.from_scaled_integer(False, (1 << (p - 1)) + 1, 1, nl)  # 2^P + 2 exactly
        half = M.round_fraction(M.Fraction(1, 1), nl)
        a = from_mp([big, half], nl)
        b = from_mp([half, M.round_fraction(M.Fraction(-1), nl)], nl)
        ok, out = gpu.rns_gemm("N", "N", 1, 1, 2, a, 1, b, 2)  # 2^P + 2 - 1 = 2^P + 1 -> tie
        exp = M.exact_dot([(big, half), (half, M.round_fraction(M.Fraction(-1), nl))], nl)
        assert ok and mp.same_bits(out, from_mp([exp], nl))[0]


def test_zero_and_empty(gpu):
    nl = 4
    a = mp.zeros(6, nl)
    b = mp.random_full(6, nl, 5, np.random.default_rng(1))
    ok, out = gpu.rns_gemm("N", "N", 2, 3, 2, a, 2, b, 2)
    assert ok and np.all(np.abs(out["kind"]) == 2)
    ok, out = gpu.rns_gemm("N", "N", 2, 3, 0, a, 2, b, 1)
    assert ok and np.all(np.abs(out["kind"]) == 2)


def test_declines_leave_output_untouched(gpu):
    nl = 4
    rng = np.random.default_rng(9)
    a = mp.random_full(40, nl, 5, rng, zero_fraction=0)
    b = mp.random_full(40, nl, 5, rng, zero_fraction=0)
    sentinel = mp.random_full(25, nl, 5, rng)
    for mutate in ("inf", "nan", "spread"):
        aa = a.copy()
        if mutate == "inf":
            aa["kind"][7] = -1
        elif mutate == "nan":
            aa["kind"][7] = 0
        else:
            aa["exponent"][3] = 4000
            aa["exponent"][4] = -200
        out = sentinel.copy()
        ok, out = gpu.rns_gemm("N", "N", 5, 5, 8, aa, 5, b, 8, out=out)
        assert not ok and np.all(mp.same_bits(out, sentinel)), mutate
    # spread exactly 4096 is admitted (rns_blas.rs:2198-2201 declines above)
    aa = a.copy()
    aa["exponent"][3] = 4096
    aa["exponent"][4] = 0
    aa["exponent"][:3] = 100
    ok, out = gpu.rns_gemm("N", "N", 5, 5, 8, aa, 5, b, 8)
    ref = oracle_gemm("N", "N", 5, 5, 8, aa, 5, b, 8)
    assert ok == (ref is not None)


def test_full_size_fixed_point_spot_check(gpu):
    # the 4096^3 fixed-point workload at 256 bits: bits of 2048 sampled
    # entries against exact dots, plus a repeat run for determinism
    rng = np.random.default_rng(2024)
    nl, m, n, k = 4, 4096, 4096, 4096
    a = mp.random_fixed(m * k, nl, 256, rng)
    b = mp.random_fixed(k * n, nl, 256, rng)
    ok, out = gpu.rns_gemm("N", "N", m, n, k, a, m, b, k)
    assert ok
    rows = rng.integers(0, m, 2048)
    cols = rng.integers(0, n, 2048)
    ref = oracle_entries("N", "N", k, a, m, b, k, rows, cols)
    assert np.all(mp.same_bits(out[rows + cols * m], ref))
    ok, out2 = gpu.rns_gemm("N", "N", m, n, k, a, m, b, k)
    assert ok and out2.tobytes() == out.tobytes()


def test_full_size_gram_spot_check(gpu):
    # the SYRK workload B^T B, B 16384 x 4096 of 256-bit fixed-point
    # entries, packed upper: the corners, the diagonal's ends and 1024
    # sampled upper entries bit for bit against exact dots over k = 16384
    rng = np.random.default_rng(16384)
    nl, k, n = 4, 16384, 4096
    print("gram", k, n, flush=True)
    b = np.empty(k * n, mp.dtype(nl))
    step = k * 256
    for c in range(0, k * n, step):  # generated in column panels to bound host memory
        b[c:c + step] = mp.random_fixed(step, nl, 256, rng)
    ok, out = gpu.rns_gemm("T", "N", n, n, k, b, k, b, k, upper_only=True)
    assert ok
    cols = np.concatenate([[0, n - 1, n - 1, 4095], rng.integers(0, n, 1024)])
    rows = np.concatenate([[0, 0, n - 1, 2048], rng.integers(0, n, 1024)])
    rows = np.minimum(rows, cols)
    ref = oracle_entries("T", "N", k, b, k, b, k, rows, cols)
    bad = ~mp.same_bits(out[cols * (cols + 1) // 2 + rows], ref)
    assert not bad.any(), f"{bad.sum()} sampled entries differ"


@pytest.mark.parametrize("m,n,k,upper", [(2100, 2300, 1100, False), (2050, 2050, 1030, True)])
def test_large_product_blocks_spot_check(gpu, m, n, k, upper):
    """Products large enough for the 256 x 256 one-digit blocks, with ragged
    edges (m, n not multiples of 256) and the packed upper Gram form: the
    corners and 700 sampled entries bit for bit against exact dots."""
    rng = np.random.default_rng(m + n + k)
    nl = 4
    print("large product", m, n, k, upper, flush=True)
    if upper:
        a = mp.random_full(k * n, nl, 30, rng)
        ok, out = gpu.rns_gemm("T", "N", n, n, k, a, k, a, k, upper_only=True)
        assert ok
        cols = np.concatenate([[0, n - 1, n - 1, 255, 256], rng.integers(0, n, 700)])
        rows = np.concatenate([[0, 0, n - 1, 255, 255], rng.integers(0, n, 700)])
        rows = np.minimum(rows, cols)  # upper entries
        ref = oracle_entries("T", "N", k, a, k, a, k, rows, cols)
        got = out[cols * (cols + 1) // 2 + rows]
    else:
        a = mp.random_full(m * k, nl, 30, rng)
        b = mp.random_full(k * n, nl, 30, rng)
        ok, out = gpu.rns_gemm("N", "N", m, n, k, a, m, b, k)
        assert ok
        rows = np.concatenate([[0, m - 1, 0, m - 1, 255, 256], rng.integers(0, m, 700)])
        cols = np.concatenate([[0, 0, n - 1, n - 1, 256, 255], rng.integers(0, n, 700)])
        ref = oracle_entries("N", "N", k, a, m, b, k, rows, cols)
        got = out[rows + cols * m]
    bad = ~mp.same_bits(got, ref)
    assert not bad.any(), f"{bad.sum()} sampled entries differ"


def test_grouped_blocks_upper(gpu):
    # xgemm_blocks_upper_exact: uneven blocks on sorted global columns, the
    # exact sum over all blocks rounded once (rns_blas.rs:2671-2922)
    import mpexact as M
    from helpers import to_mp, from_mp
    rng = np.random.default_rng(17)
    nl, m = 4, 40
    blocks = []
    for rows, ncols in ((7, 12), (30, 40), (3, 5)):
        cols = np.sort(rng.choice(m, ncols, replace=False)).astype(np.uint64)
        y = mp.random_full(rows * ncols, nl, 20, rng)
        d = mp.random_full(rows, nl, 10, rng)
        blocks.append((y, d, cols))
    out = mp.zeros(m * m, nl)
    assert gpu.xgemm_blocks_upper_exact(m, blocks, out)
    # oracle: exact sums with z = RN(d y)
    for i in range(m):
        for j in range(i, m, 7):
            pairs = []
            for y, d, cols in blocks:
                cl = list(cols)
                if i in cl and j in cl:
                    ci, cj = cl.index(i), cl.index(j)
                    rows = d.size
                    for r in range(rows):
                        yi = to_mp(y[r + ci * rows: r + ci * rows + 1])[0]
                        yj = to_mp(y[r + cj * rows: r + cj * rows + 1])[0]
                        dr = to_mp(d[r:r + 1])[0]
                        pairs.append((M.mul(dr, yi), yj))
            ref = M.exact_dot(pairs, nl)
            assert mp.same_bits(out[i + j * m: i + j * m + 1], from_mp([ref], nl))[0], (i, j)
    # the strict lower triangle is left as given
    sentinel = mp.from_float64(np.full(m * m, -0.375), nl)
    kept = sentinel.copy()
    assert gpu.xgemm_blocks_upper_exact(m, blocks, kept)
    for j in range(m):
        for i in range(m):
            e = i + j * m
            same = kept[e].tobytes() == (out[e] if i <= j else sentinel[e]).tobytes()
            assert same, (i, j)
    # the device-pointer form gives the same bits
    import torch
    dev = lambda rec: torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8).copy()).cuda()
    keep = []
    arr = (gpu.ExactBlock * len(blocks))()
    for b, (y, d, cols) in enumerate(blocks):
        ty, td, tc = dev(y), dev(d), np.ascontiguousarray(cols, np.uint64)
        keep += [ty, td, tc]
        arr[b] = gpu.ExactBlock(d.size, cols.size, ty.data_ptr(), td.data_ptr(), tc.ctypes.data)
    tout = dev(mp.zeros(m * m, nl))
    torch.cuda.synchronize()
    L = gpu.lib()
    L.sdpx_gpu_rns_gemm_blocks_upper_dev.restype = ctypes.c_int32
    gpu.check(L.sdpx_gpu_rns_gemm_blocks_upper_dev(nl, m, len(blocks), arr, ctypes.c_void_p(tout.data_ptr())))
    got = tout.cpu().numpy().view(mp.dtype(nl))
    for j in range(m):
        assert got[j * m: j * m + j + 1].tobytes() == out[j * m: j * m + j + 1].tobytes(), j


@pytest.mark.parametrize("bits", [1, 2, 3, 8, 13, 24, 31, 40, 53])
def test_narrow_operands_every_small_plan(gpu, bits):
    """Operands of a few significant bits (small integers scaled by powers
    of two) give integer products of a few dozen bits, so the plans have
    one, two or three moduli words: the CRT's quotient estimate must scale
    M and V alike for every word count (a one-word plan once lifted 1 to
    1 - (2^31 - 1) M).  Includes the 2 x 2 x 2 product of the analytic
    sampled block (basis [1, 1; 1, -1])."""
    rng = np.random.default_rng(bits)
    for nl in (2, 4, 8):
        for m, n, k in ((2, 2, 2), (5, 3, 7), (64, 65, 33), (130, 129, 300)):
            for ta, tb in (("N", "N"), ("T", "N")):
                ar, ac = (m, k) if ta == "N" else (k, m)
                br, bc = (k, n) if tb == "N" else (n, k)
                lim = 2 ** (bits - 1)
                av = rng.integers(-lim, lim + 1, ar * ac) * np.ldexp(1.0, rng.integers(-4, 5, ar * ac))
                bv = rng.integers(-lim, lim + 1, br * bc) * np.ldexp(1.0, rng.integers(-4, 5, br * bc))
                a, b = mp.from_float64(av, nl), mp.from_float64(bv, nl)
                ok, out = gpu.rns_gemm(ta, tb, m, n, k, a, ar, b, br)
                ref = oracle_gemm(ta, tb, m, n, k, a, ar, b, br)
                assert ok
                assert np.all(mp.same_bits(out, ref)), (bits, nl, m, n, k, ta, tb)
                if m == n:
                    ok, up = gpu.rns_gemm(ta, tb, m, n, k, a, ar, b, br, upper_only=True)
                    refp = oracle_gemm(ta, tb, m, n, k, a, ar, b, br, upper=True)
                    assert ok and np.all(mp.same_bits(up, refp))
    q = mp.from_float64([1.0, 1.0, 1.0, -1.0], 2)
    eye = mp.from_float64([1.0, 0.0, 0.0, 1.0], 2)
    ok, v = gpu.rns_gemm("N", "N", 2, 2, 2, eye, 2, q, 2)
    assert ok and np.all(mp.same_bits(v, q))
