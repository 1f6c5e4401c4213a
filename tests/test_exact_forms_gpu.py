"""The exact multilinear forms on the GPU (sdpx_gpu_diag_congruence_upper_exact,
_congruence_exact, _svec_quadratic_exact, _symmetric_bilinear_exact and
their _dev forms), bit for bit against the oracle (oracle/mpexact.py) on the
known answers of tests/exact_forms_cases.py and on random operands of the
reference's test distribution."""
from fractions import Fraction

import numpy as np
import pytest

import mpexact
import pins
from exact_forms_cases import random_forms, rounding_once_case, sqrt2_of, weight_case
from helpers import from_mp, to_mp
from sdpx_gpu import mp

pytestmark = pytest.mark.gpu

SENTINEL = Fraction(-7, 3 * 2 ** 40)


def keys(vals):
    return [None if v is None else v.key() for v in vals]


def sentinel_records(count, nl):
    return pins.enc([pins.rn(SENTINEL, 64 * nl)] * count, nl)


def check_upper(c, want, m, nl):
    """c (m x m records) holds want's entries where want has them and the
    untouched sentinel elsewhere."""
    got = to_mp(c)
    sent = pins.enc([pins.rn(SENTINEL, 64 * nl)], nl)
    for e, w in enumerate(want):
        if w is None:
            assert c[e].tobytes() == sent[0].tobytes(), e
        else:
            assert got[e].key() == w.key(), (e % m, e // m)


@pytest.mark.parametrize("nl", [2, 4, 8])
def test_rounding_once_known_answer_gpu(gpu, nl):
    p = 64 * nl
    a, d, once, _ = rounding_once_case(p)
    av, one = pins.enc([a, 1], nl), Fraction(1)
    c = mp.zeros(1, nl)
    assert gpu.diag_congruence_upper_exact(1, 2, av, pins.enc([d, -one], nl), c)
    assert pins.dec(c) == [once]
    if nl == 8:
        return
    for ta, lda in (("N", 1), ("T", 2)):
        c = mp.zeros(1, nl)
        assert gpu.xcongruence_exact(ta, 1, 2, av, lda, pins.enc([d, 0, 0, -one], nl), 2, c, False)
        assert pins.dec(c) == [once]
    out = mp.zeros(1, nl)
    s = pins.enc([sqrt2_of(nl)], nl)
    assert gpu.xsvec_quadratic_exact(2, 1, av, False, pins.enc([d, 0, -one], nl), s, out)
    assert pins.dec(out) == [once]
    out = mp.zeros(1, nl)
    assert gpu.xsymmetric_bilinear_exact(2, 1, av, pins.enc([d, 0, -one], nl), [(0, 0)], out)
    assert pins.dec(out) == [once]


@pytest.mark.parametrize("nl", [2, 4])
def test_svec_weights_gpu(gpu, nl):
    q, x, s, expected = weight_case(nl)
    for qv, absq, want in ((q, False, expected), ([-1, 1], True, expected),
                           ([-1, 1], False, pins.rn(1 - s * s, 64 * nl))):
        out = mp.zeros(1, nl)
        assert gpu.xsvec_quadratic_exact(2, 1, pins.enc(qv, nl), absq, pins.enc(x, nl), pins.enc([s], nl), out)
        assert pins.dec(out) == [want]


@pytest.mark.parametrize("nl", [2, 4, 8])
@pytest.mark.parametrize("spread", [6, 60])
def test_diag_congruence_random_gpu(gpu, nl, spread):
    print("diag_congruence", nl, spread, flush=True)
    m, k, a, d = random_forms(nl, spread, 100 + spread, m=36, k=70)["diag"]
    c = sentinel_records(m * m, nl)
    assert gpu.diag_congruence_upper_exact(m, k, from_mp(a, nl), from_mp(d, nl), c)
    check_upper(c, mpexact.diag_congruence_exact(m, k, a, d), m, nl)


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("upper", [False, True])
def test_congruence_random_gpu(gpu, nl, ta, upper):
    print("congruence", nl, ta, upper, flush=True)
    m, k, a, x = random_forms(nl, 20, 7, m=22, k=37)["cong_n" if ta == "N" else "cong_t"]
    lda = m if ta == "N" else k
    c = sentinel_records(m * m, nl)
    assert gpu.xcongruence_exact(ta, m, k, from_mp(a, nl), lda, from_mp(x, nl), k, c, upper)
    check_upper(c, mpexact.congruence_exact(ta, m, k, a, lda, x, k, upper), m, nl)


@pytest.mark.parametrize("nl", [2, 4])
@pytest.mark.parametrize("abs_q", [False, True])
@pytest.mark.parametrize("weight", ["sqrt2", "two"])
def test_svec_quadratic_random_gpu(gpu, nl, abs_q, weight):
    print("svec_quadratic", nl, abs_q, weight, flush=True)
    h, kmax, q, x = random_forms(nl, 20, 21, h=19, kmax=23)["quad"]
    s = sqrt2_of(nl) if weight == "sqrt2" else Fraction(2)
    sv = pins.enc([s], nl)
    out = mp.zeros(kmax, nl)
    assert gpu.xsvec_quadratic_exact(h, kmax, from_mp(q, nl), abs_q, from_mp(x, nl), sv, out)
    want = mpexact.svec_quadratic_exact(h, kmax, q, abs_q, x, to_mp(sv)[0])
    assert keys(to_mp(out)) == keys(want)


@pytest.mark.parametrize("nl", [2, 4])
def test_symmetric_bilinear_random_gpu(gpu, nl):
    print("symmetric_bilinear", nl, flush=True)
    h, cols, q, x, pairs = random_forms(nl, 20, 33, h=17, cols=13, npairs=40)["bil"]
    out = mp.zeros(len(pairs), nl)
    assert gpu.xsymmetric_bilinear_exact(h, cols, from_mp(q, nl), from_mp(x, nl), pairs, out)
    assert keys(to_mp(out)) == keys(mpexact.symmetric_bilinear_exact(h, cols, q, x, pairs))


def test_expansion_forms_at_512_bits_match_or_decline(gpu):
    """At 512 bits the exact intermediate of congruence needs more than the
    residue encoder's 1536-bit operand width, so the form declines (the
    caller takes its own path); where it does not, the bits match."""
    nl = 8
    m, k, a, x = random_forms(nl, 4, 5, m=6, k=9)["cong_n"]
    c = mp.zeros(m * m, nl)
    if gpu.xcongruence_exact("N", m, k, from_mp(a, nl), m, from_mp(x, nl), k, c, False):
        assert keys(to_mp(c)) == keys(mpexact.congruence_exact("N", m, k, a, m, x, k))


@pytest.mark.parametrize("nl", [2, 4])
def test_zero_operands_and_empty_sums_gpu(gpu, nl):
    m, k = 5, 4
    zeros = mp.zeros(k * m, nl)
    d = pins.enc([1, 2, 3, 4], nl)
    c = sentinel_records(m * m, nl)
    assert gpu.diag_congruence_upper_exact(m, k, zeros, d, c)
    check_upper(c, [mpexact.zero(nl) if i <= j else None for j in range(m) for i in range(m)], m, nl)
    c = sentinel_records(m * m, nl)
    assert gpu.xcongruence_exact("N", m, 0, None, m, None, 1, c, False)
    assert pins.dec(c) == [0] * (m * m)
    out = sentinel_records(3, nl)
    assert gpu.xsvec_quadratic_exact(2, 3, mp.zeros(6, nl), False, pins.enc([1, 2, 3], nl),
                                     pins.enc([sqrt2_of(nl)], nl), out)
    assert pins.dec(out) == [0, 0, 0]
    out = mp.zeros(1, nl)
    assert not gpu.xsymmetric_bilinear_exact(0, 1, None, None, [(0, 0)], out)


def test_dev_entries_match_host_entries(gpu):
    import torch

    nl = 2
    lib = gpu.lib()
    dev = lambda rec: torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8).copy()).cuda()
    host = lambda t, count: t.cpu().numpy().view(mp.dtype(nl))[:count].copy()
    f = random_forms(nl, 20, 44, m=12, k=15, h=9, kmax=7, cols=6, npairs=9)
    m, k, a, d = f["diag"]
    ra, rd = from_mp(a, nl), from_mp(d, nl)
    want = sentinel_records(m * m, nl)
    assert gpu.diag_congruence_upper_exact(m, k, ra, rd, want)
    da, dd, dc = dev(ra), dev(rd), dev(sentinel_records(m * m, nl))
    torch.cuda.synchronize()
    gpu.check(lib.sdpx_gpu_diag_congruence_upper_exact_dev(nl, m, k, da.data_ptr(), dd.data_ptr(), dc.data_ptr(), 1))
    assert host(dc, m * m).tobytes() == want.tobytes()
    m, k, a, x = f["cong_t"]
    ra, rx = from_mp(a, nl), from_mp(x, nl)
    want = mp.zeros(m * m, nl)
    assert gpu.xcongruence_exact("T", m, k, ra, k, rx, k, want, False)
    da, dx, dc = dev(ra), dev(rx), dev(mp.zeros(m * m, nl))
    torch.cuda.synchronize()
    gpu.check(lib.sdpx_gpu_congruence_exact_dev(nl, ord("T"), m, k, da.data_ptr(), k, dx.data_ptr(), k,
                                                dc.data_ptr(), 0))
    assert host(dc, m * m).tobytes() == want.tobytes()
    h, kmax, q, x = f["quad"]
    rq, rx, rs = from_mp(q, nl), from_mp(x, nl), pins.enc([sqrt2_of(nl)], nl)
    want = mp.zeros(kmax, nl)
    assert gpu.xsvec_quadratic_exact(h, kmax, rq, True, rx, rs, want)
    dq, dx, ds, do = dev(rq), dev(rx), dev(rs), dev(mp.zeros(kmax, nl))
    torch.cuda.synchronize()
    gpu.check(lib.sdpx_gpu_svec_quadratic_exact_dev(nl, h, kmax, dq.data_ptr(), 1, dx.data_ptr(), ds.data_ptr(),
                                                    do.data_ptr()))
    assert host(do, kmax).tobytes() == want.tobytes()
    h, cols, q, x, pairs = f["bil"]
    rq, rx = from_mp(q, nl), from_mp(x, nl)
    want = mp.zeros(len(pairs), nl)
    assert gpu.xsymmetric_bilinear_exact(h, cols, rq, rx, pairs, want)
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1))
    dq, dx, do = dev(rq), dev(rx), dev(mp.zeros(len(pairs), nl))
    torch.cuda.synchronize()
    gpu.check(lib.sdpx_gpu_symmetric_bilinear_exact_dev(nl, h, cols, dq.data_ptr(), dx.data_ptr(), len(pairs),
                                                        pr.ctypes.data, do.data_ptr()))
    assert host(do, len(pairs)).tobytes() == want.tobytes()
