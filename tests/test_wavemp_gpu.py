"""The wave-distributed arithmetic of the panel kernels
(sdpx.rs_amd/csrc/wavemp.h) agrees bit for bit with the register routines
of mpdev.h (themselves pinned to the host library and the oracle by
test_mpdev.py) on the GPU: sums and differences across every exponent gap
class (cancellation, gaps around the word size and around the precision,
exact ties), products (with overflow to infinity and underflow to zero),
reciprocals, sign tests, the rounding of exact-dot windows with and
without an added exact product, and the fused multiply-subtract
RN(c - a b) over every gap class between c and the product.  Each case runs on one wave in both forms;
tests/native/libwavemp_selftest.so is built by build()."""
import ctypes
import os

import numpy as np
import pytest

from sdpx_gpu import mp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libwavemp_selftest.so")
EMAX, EMIN = (1 << 30) - 1, 1 - (1 << 30)


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libwavemp_selftest.so is not built (run build())")
    lib = ctypes.CDLL(LIB)
    lib.wmp_selftest.restype = ctypes.c_int
    lib.wmp_selftest.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int] + \
        [ctypes.c_void_p] * 2
    return lib


def _run(lib, nl, op, a, b, T=None, bo=None):
    n = a.size
    w = 2 * nl
    if T is None:
        T = np.zeros((n, 64), np.int64)
    if bo is None:
        bo = np.zeros((n, 2), np.int64)
    A = np.ascontiguousarray(a).view(np.uint32).reshape(n, w + 2)
    B = np.ascontiguousarray(b).view(np.uint32).reshape(n, w + 2)
    S = np.zeros((n, w + 2), np.uint32)
    O = np.zeros((n, w + 2), np.uint32)
    T = np.ascontiguousarray(T, np.int64)
    bo = np.ascontiguousarray(bo, np.int64)
    rc = lib.wmp_selftest(nl, op, A.ctypes.data, B.ctypes.data, T.ctypes.data, bo.ctypes.data, n,
                          S.ctypes.data, O.ctypes.data)
    assert rc == 0, f"HIP error {rc}"
    return S, O


def _pow2(count, nl, exps, neg):
    r = mp.zeros(count, nl)
    r["limbs"][:, -1] = np.uint64(1 << 63)
    r["kind"] = np.where(neg, -mp.REGULAR_KIND, mp.REGULAR_KIND)
    r["exponent"] = exps
    return r


def _operands(nl, rng, count=3000):
    """Pairs (a, b) covering every exponent-gap class of a sum."""
    P = 64 * nl
    a = mp.random_full(count, nl, 40, rng)
    b = mp.random_full(count, nl, 40, rng)
    # gaps around words and around the precision, both orders
    gaps = np.array([0, 1, 2, 31, 32, 33, 63, 64, 65, P - 33, P - 1, P, P + 1, P + 2, P + 3, P + 64])
    g = rng.choice(gaps, size=count // 3)
    idx = rng.choice(count, size=count // 3, replace=False)
    reg = np.abs(a["kind"][idx]) == mp.REGULAR_KIND
    b["limbs"][idx] = mp.random_full(idx.size, nl, 0, rng, zero_fraction=0)["limbs"]
    b["kind"][idx] = np.where(rng.integers(0, 2, idx.size) > 0, 3, -3)
    b["exponent"][idx] = np.where(reg, a["exponent"][idx] - g, 0)
    sw = rng.integers(0, 2, idx.size).astype(bool)
    tmp = a[idx[sw]].copy()
    a[idx[sw]] = b[idx[sw]]
    b[idx[sw]] = tmp
    # near cancellation: b = -a with the low words changed; exact negatives
    k = count // 10
    idx = rng.choice(count, size=k, replace=False)
    b[idx] = a[idx]
    b["kind"][idx] = -a["kind"][idx]
    half = idx[: k // 2]
    b["limbs"][half, 0] ^= rng.integers(0, 2**20, half.size, dtype=np.uint64)
    # ties: a power-of-two addend half an ulp (and a quarter, and one ulp) below
    k = count // 10
    idx = rng.choice(count, size=k, replace=False)
    d = rng.choice(np.array([P, P + 1, P - 1]), size=k)
    b[idx] = _pow2(k, nl, a["exponent"][idx] - d, rng.integers(0, 2, k).astype(bool))
    zero_a = np.abs(a["kind"][idx]) == mp.ZERO_KIND
    b["kind"][idx[zero_a]] = mp.ZERO_KIND
    b["limbs"][idx[zero_a]] = 0
    b["exponent"][idx[zero_a]] = 0
    # all-ones mantissas (carries through every word)
    k = count // 20
    idx = rng.choice(count, size=k, replace=False)
    a["limbs"][idx] = np.uint64(2**64 - 1)
    a["kind"][idx] = 3
    return a, b


def _check(S, O, what, kinds=(mp.REGULAR_KIND,)):
    bad = np.nonzero(np.any(S != O, axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} cases differ, first {bad[:5]}: {S[bad[0]]} vs {O[bad[0]]}"
    got = np.abs(S[:, -2].view(np.int32))
    for kind in kinds:  # the cases reach every class of result named
        assert np.any(got == kind), f"{what}: no result of kind {kind}"


@pytest.mark.gpu
@pytest.mark.parametrize("nl", [2, 4, 8])
def test_wave_add_sub_compare(gpu, nl):
    lib = _lib()
    rng = np.random.default_rng(100 + nl)
    a, b = _operands(nl, rng)
    for op, what in ((0, "a + b"), (1, "a - b"), (4, "sign(a - b)")):
        S, O = _run(lib, nl, op, a, b)
        _check(S, O, what, (mp.REGULAR_KIND, mp.ZERO_KIND) if op < 4 else (mp.ZERO_KIND,))
        if op == 4:
            assert set(S[:, 0].tolist()) == {0, 1, 0xFFFFFFFF}


@pytest.mark.gpu
@pytest.mark.parametrize("nl", [2, 4, 8])
def test_wave_mul_recip(gpu, nl):
    lib = _lib()
    rng = np.random.default_rng(200 + nl)
    n = 3000
    a = mp.random_full(n, nl, 200, rng)
    b = mp.random_full(n, nl, 200, rng)
    k = n // 10
    a["limbs"][:k] = np.uint64(2**64 - 1)
    a["kind"][:k] = 3
    b["limbs"][k:2 * k] = np.uint64(2**64 - 1)
    b["kind"][k:2 * k] = -3
    a[2 * k:3 * k] = _pow2(k, nl, rng.integers(-50, 50, k), rng.integers(0, 2, k).astype(bool))
    # one word below the top set: 0x8000..0001-like mantissas
    a["limbs"][3 * k:4 * k] = 0
    a["limbs"][3 * k:4 * k, -1] = np.uint64(1 << 63)
    a["limbs"][3 * k:4 * k, 0] = rng.integers(1, 2**64, k, dtype=np.uint64)
    a["kind"][3 * k:4 * k] = 3
    # exponents at the ends of the range: overflow to inf, underflow to zero
    e = rng.choice(np.array([EMAX, EMAX - 1, EMIN, EMIN + 1, EMAX // 2 + 1, EMIN // 2 - 1]), size=(k, 2))
    reg = (np.abs(a["kind"][4 * k:5 * k]) == 3) & (np.abs(b["kind"][4 * k:5 * k]) == 3)
    a["exponent"][4 * k:5 * k] = np.where(reg, e[:, 0], a["exponent"][4 * k:5 * k])
    b["exponent"][4 * k:5 * k] = np.where(reg, e[:, 1], b["exponent"][4 * k:5 * k])
    S, O = _run(lib, nl, 2, a, b)
    _check(S, O, "a * b", (mp.REGULAR_KIND, mp.ZERO_KIND, mp.INF_KIND))
    nz = np.abs(a["kind"]) == 3
    S, O = _run(lib, nl, 3, a[nz], b[nz])
    _check(S, O, "1 / a", (mp.REGULAR_KIND, mp.INF_KIND))
    zs = mp.zeros(4, nl)
    zs["kind"][1::2] = -mp.ZERO_KIND
    S, O = _run(lib, nl, 3, zs, zs)
    _check(S, O, "1 / 0", (mp.INF_KIND,))


@pytest.mark.gpu
@pytest.mark.parametrize("nl", [2, 4, 8])
def test_wave_window_rounding(gpu, nl):
    lib = _lib()
    rng = np.random.default_rng(300 + nl)
    W = 2 * nl
    SL = 2 * W + 20
    n = 3000
    T = np.zeros((n, 64), np.int64)
    top = rng.integers(1, SL - 1, n)
    mag = rng.choice(np.array([8, 31, 33, 40, 62]), size=n)
    for c in range(n):
        lim = np.int64(1) << np.int64(mag[c])
        T[c, : top[c]] = rng.integers(-lim, lim, top[c])
    T[: n // 10, :] = 0                       # empty windows
    T[n // 10: n // 5, :] = 0                 # a single word
    T[np.arange(n // 10, n // 5), top[n // 10: n // 5]] = rng.integers(-2**31, 2**31, n // 10)
    bo = np.stack([rng.integers(-3000, 3000, n), rng.integers(0, 32 * (SL - 2 - 2 * W), n)], axis=1)
    # carry and borrow chains through many slots: runs of all-ones or zero
    # words with a +1 / -1 entering at the bottom, and a small top slot
    m32 = np.int64(0xFFFFFFFF)
    for c in range(n // 3, n // 3 + 400):
        run = int(rng.integers(2, SL - 2))
        T[c, :] = 0
        if c % 2:
            T[c, :run] = m32
            T[c, 0] = m32 + rng.integers(1, 3)          # +1 / +2 ripples up the ones
        else:
            T[c, 0] = -rng.integers(1, 3)               # borrows through the zeros
            T[c, run] = rng.integers(1, 2**31)
        T[c, SL - 1] = rng.integers(-2, 3)              # the top slot keeps the sign
    z = mp.zeros(n, nl)
    S, O = _run(lib, nl, 5, z, z, T, bo)
    _check(S, O, "window", (mp.REGULAR_KIND, mp.ZERO_KIND))
    a = mp.random_full(n, nl, 60, rng)
    b = mp.random_full(n, nl, 60, rng)
    T[n // 5: n // 3, :] = 0                  # the product alone (the chains above stay)
    S, O = _run(lib, nl, 6, a, b, T, bo)
    _check(S, O, "window + a * b")


@pytest.mark.gpu
@pytest.mark.parametrize("nl", [2, 4, 8])
def test_wave_fms(gpu, nl):
    lib = _lib()
    rng = np.random.default_rng(400 + nl)
    P = 64 * nl
    # products equal to a (b = +-1): every gap class of a difference, with
    # cancellation, ties and all-ones carries
    c1, a1 = _operands(nl, rng)
    b1 = _pow2(a1.size, nl, np.ones(a1.size, np.int64), rng.integers(0, 2, a1.size).astype(bool))
    # full 2W-word products with c placed at each gap class around the
    # product's exponent a.e + b.e (both signs, the far cases' edges)
    n = 4000
    a2 = mp.random_full(n, nl, 30, rng, zero_fraction=0)
    b2 = mp.random_full(n, nl, 30, rng, zero_fraction=0)
    c2 = mp.random_full(n, nl, 0, rng, zero_fraction=0)
    gaps = np.array([-3 * P - 64, -2 * P - 4, -2 * P - 3, -2 * P - 2, -2 * P - 1, -2 * P, -P - 1, -P, -P + 1,
                     -64, -33, -32, -1, 0, 1, 2, 31, 32, 33, P - 1, P, P + 1, P + 2, P + 3, 2 * P, 3 * P + 64])
    reg = (np.abs(a2["kind"]) == 3) & (np.abs(b2["kind"]) == 3) & (np.abs(c2["kind"]) == 3)
    c2["exponent"] = np.where(reg, a2["exponent"] + b2["exponent"] + rng.choice(gaps, n), c2["exponent"])
    # zeros in each operand
    k = n // 20
    for arr, sl in ((a2, slice(0, k)), (b2, slice(k, 2 * k)), (c2, slice(2 * k, 3 * k))):
        arr["limbs"][sl] = 0
        arr["kind"][sl] = mp.ZERO_KIND
        arr["exponent"][sl] = 0
    a = np.concatenate([a1, a2])
    b = np.concatenate([b1, b2])
    c = np.concatenate([c1, c2])
    T = np.zeros((a.size, 64), np.int64)
    T.view(np.uint32)[:, : 2 * nl + 2] = np.ascontiguousarray(c).view(np.uint32).reshape(a.size, 2 * nl + 2)
    S, O = _run(lib, nl, 7, a, b, T=T)
    _check(S, O, "c - a * b", (mp.REGULAR_KIND, mp.ZERO_KIND))
