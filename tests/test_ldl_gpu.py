"""Multiple-precision LDL^T with every exact product of the recursion on
the GPU, bitwise against the C oracle (itself pinned in test_ldl.py)."""
import numpy as np
import pytest

import sdpx_gpu
from helpers import oracle_ldl
from sdpx_gpu import mp
from test_ldl import reference_input

pytestmark = pytest.mark.gpu


def test_reference_input_gpu_products(gpu):
    n, nl = 300, 2
    A, signs, reg, _, _ = reference_input(n, nl)
    ok, L, dinv, count = oracle_ldl(A, n, signs, reg)
    okg, Lg, dg = sdpx_gpu.mp_ldl(A, n, signs, reg)
    st = sdpx_gpu.mp_ldl_stats()
    assert st["gpu_products"] >= 1 and st["regularized"] == count and st["declined"] == 0
    assert okg and np.all(mp.same_bits(Lg, L)) and np.all(mp.same_bits(dg, dinv))


def spread_spd(n, rng, lo_exp=-30):
    """Q diag(s) Q^T with s log-spaced over [10^lo_exp, 1] (float64 data)."""
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(lo_exp, 0, n)
    return (q * s) @ q.T


@pytest.mark.parametrize("nl,n,kind", [(4, 600, "gram"), (8, 520, "spread"), (8, 1030, "spread")])
def test_random_gpu_products(gpu, nl, n, kind):
    rng = np.random.default_rng(n + nl)
    print("ldl", nl, n, kind, flush=True)
    if kind == "gram":
        g = rng.standard_normal((n, n))
        s = g @ g.T / n + 0.05 * np.eye(n)
    else:
        s = spread_spd(n, rng, -12)
        s = (s + s.T) / 2
    A = mp.from_float64(s.T.reshape(-1), nl)
    ok, L, dinv, _ = oracle_ldl(A, n, threads=16)
    okg, Lg, dg = sdpx_gpu.mp_ldl(A, n)
    st = sdpx_gpu.mp_ldl_stats()
    assert st["gpu_products"] >= 2 and st["declined"] == 0
    assert ok and okg
    bad = ~mp.same_bits(Lg, L)
    assert not bad.any(), f"{bad.sum()} factor entries differ"
    assert np.all(mp.same_bits(dg, dinv))


@pytest.mark.parametrize("n", [300, 100])
def test_failed_pivot_reported(gpu, n):
    """A zero row and column make one pivot exactly zero: the device path
    (recursive at 300, one ldl_dots block at 100) reports the failed pivot
    as the oracle does, after queueing the whole factorization."""
    nl = 4
    rng = np.random.default_rng(n)
    g = rng.standard_normal((n, n))
    s = g @ g.T / n + np.eye(n)
    z = n // 2 + 3
    s[z, :] = 0.0
    s[:, z] = 0.0
    A = mp.from_float64(s.T.reshape(-1), nl)
    ok, _, _, _ = oracle_ldl(A, n, threads=16)
    okg, _, _ = sdpx_gpu.mp_ldl(A, n)
    assert not ok and not okg and sdpx_gpu.mp_ldl_stats()["declined"] == 0


_spread_case = {}


@pytest.mark.parametrize("room", [-100000, 0, 128])
def test_trsm_right_looking_rows_and_fallback(gpu, room):
    """The triangular-solve base's right-looking pass: with a negative room
    every row goes to the left-looking kernel, with no room below the first
    term some rows do, and by default (128 bits) few or none; the factor is
    the oracle's bit for bit in each case."""
    nl, n = 4, 520
    if not _spread_case:
        rng = np.random.default_rng(77)
        s = spread_spd(n, rng, -12)
        A = mp.from_float64(((s + s.T) / 2).T.reshape(-1), nl)
        _spread_case["A"] = A
        _spread_case["ref"] = oracle_ldl(A, n, threads=16)
    A = _spread_case["A"]
    ok, L, dinv, _ = _spread_case["ref"]
    print("trsm room", room, flush=True)
    sdpx_gpu.set_tuning("trsm_rl_room", room)
    try:
        okg, Lg, dg = sdpx_gpu.mp_ldl(A, n)
        declined = sdpx_gpu.mp_ldl_stats()["declined"]
    finally:
        sdpx_gpu.set_tuning("trsm_rl_room", 128)
    assert ok and okg and declined == 0
    assert np.all(mp.same_bits(Lg, L)) and np.all(mp.same_bits(dg, dinv))


def widen(rec, nl_to):
    """Exact widening of records to 64*nl_to bits (zero low limbs)."""
    nl = mp.limbs_of(rec)
    out = mp.zeros(rec.size, nl_to)
    out["limbs"][:, nl_to - nl:] = rec["limbs"]
    out["kind"] = rec["kind"]
    out["exponent"] = rec["exponent"]
    return out


def test_solves_residual_vs_1024_bit_oracle(gpu):
    # L D L^T x = b at 512 bits for 16 Gaussian right-hand sides; the
    # residual b - A x is evaluated exactly at 1024 bits (C oracle exact
    # dots on exactly widened values)
    from helpers import oracle_gemm
    rng = np.random.default_rng(8)
    nl, n, nrhs = 8, 400, 16
    s = spread_spd(n, rng, -30)
    s = (s + s.T) / 2
    A = mp.from_float64(s.T.reshape(-1), nl)
    okg, L, dinv = sdpx_gpu.mp_ldl(A, n)
    assert okg
    b = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    x = sdpx_gpu.mp_ldl_solve(L, dinv, n, b)
    Aw, xw, bw = widen(A, 16), widen(x, 16), widen(b, 16)
    ax = oracle_gemm("N", "N", n, nrhs, n, Aw, n, xw, n, threads=16)
    r = mp.to_float64(ax) - mp.to_float64(bw)  # |r| tiny: float64 views of 1024-bit values
    scale = np.abs(s).max() * np.abs(mp.to_float64(x)).max() + np.abs(mp.to_float64(b)).max()
    # Ax is exact to 1024 bits, so the float64 difference can only see
    # residuals above 2^-53 relative; demand far less: every entry of
    # b - Ax at 1024 bits below 2^-400 relative
    from helpers import to_mp
    import mpexact as M
    worst = 0
    axm, bm = to_mp(ax), to_mp(bw)
    for t in range(0, n * nrhs, 37):
        d = abs(axm[t].to_fraction() - bm[t].to_fraction())
        worst = max(worst, float(d / M.Fraction(scale)) if d else 0.0)
    assert worst < 2.0 ** -400, worst
    assert np.abs(r).max() <= 1e-12 * scale


def test_device_pointer_solve_matches_host_entry(gpu):
    # sdpx_gpu_mp_ldl_solve_dev on HBM-resident factor and right-hand sides
    # gives the host-pointer entry's bits
    import torch
    rng = np.random.default_rng(21)
    nl, n, nrhs = 4, 300, 5
    s = spread_spd(n, rng, -20)
    s = (s + s.T) / 2
    A = mp.from_float64(s.T.reshape(-1), nl)
    okg, L, dinv = sdpx_gpu.mp_ldl(A, n)
    assert okg
    b = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    x_host = sdpx_gpu.mp_ldl_solve(L, dinv, n, b)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to("cuda")

    dl, dd, db = dev(L), dev(dinv), dev(b)
    sdpx_gpu.check(sdpx_gpu.lib().sdpx_gpu_mp_ldl_solve_dev(nl, n, dl.data_ptr(), n, dd.data_ptr(), nrhs,
                                                            db.data_ptr(), n))
    torch.cuda.synchronize()
    x_dev = db.cpu().numpy().view(mp.dtype(nl))
    assert np.all(mp.same_bits(x_dev, x_host))


@pytest.mark.parametrize("one_launch", [0, 1])
def test_column_launches_and_one_launch_agree(gpu, one_launch):
    """ldl_dots blocks as one launch of row workgroups ordered by flags, or
    as one launch per column: both give the oracle's bits (recursive order
    600 at 256 bits: base blocks of 75 and 150 columns, plus a dots-only
    order 200 at 512 bits)."""
    rng = np.random.default_rng(91)
    print("ldl one_launch", one_launch, flush=True)
    sdpx_gpu.set_tuning("ldl_one_launch", one_launch)
    try:
        for nl, n, flags in ((4, 600, 0), (8, 200, sdpx_gpu.LDL_DOTS)):
            s = spread_spd(n, rng, -12)
            A = mp.from_float64(((s + s.T) / 2).T.reshape(-1), nl)
            ok, L, dinv, _ = oracle_ldl(A, n, recursive=not flags, threads=16)
            okg, Lg, dg = sdpx_gpu.mp_ldl(A, n, flags=flags)
            assert sdpx_gpu.mp_ldl_stats()["declined"] == 0
            assert ok and okg
            assert np.all(mp.same_bits(Lg, L)) and np.all(mp.same_bits(dg, dinv))
    finally:
        sdpx_gpu.set_tuning("ldl_one_launch", 1)


def test_order_8192_factor_properties(gpu):
    """The full-size dense factor: LDL^T at 512 bits of A = Q diag(s) Q^T of
    order 8192, s = 2^-100 ... 1 in powers of two and Q a seeded orthogonal
    matrix (A formed exactly rounded by one exact product), device-resident.
    Size-independent properties: every pivot positive; sum log2 D_i equals
    log2 det A = sum log2 s_i (Q^T Q = I to 2^-40, the 512-bit rounding of
    A moves det A by far less); a second factor of the same input gives the
    same bits; 16 Gaussian right-hand sides solve to a residual b - A x,
    with A x exactly rounded at 1024 bits, below 2^-400 of max |b|."""
    import ctypes

    import torch

    nl, n, nrhs = 8, 8192, 16
    print("order", n, "LDL^T at 512 bits", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(11)
    g = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    q = torch.linalg.qr(g)[0].cpu().numpy()
    del g
    assert np.abs(q.T @ q[:, :64] - np.eye(n)[:, :64]).max() < 2.0 ** -40
    e = np.round(np.linspace(-100, 0, n))
    s = 2.0 ** e

    def dev(rec):
        return torch.from_numpy(rec.view(np.uint8).reshape(-1)).to("cuda")

    vb = mp.dtype(nl).itemsize
    dq = dev(mp.from_float64(q.T.reshape(-1), nl))         # Q column-major
    dbq = dev(mp.from_float64((q * s).T.reshape(-1), nl))  # Q diag(s), exact
    da = torch.empty(n * n * vb, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    assert sdpx_gpu.rns_gemm_dev("N", "T", n, n, n, dbq.data_ptr(), n, dq.data_ptr(), n, nl, False,
                                 da.data_ptr(), n * n) == sdpx_gpu.OK
    del dq, dbq
    L = sdpx_gpu.lib()
    work = da.clone()
    dinv = torch.empty(n * vb, dtype=torch.uint8, device="cuda")
    ok = ctypes.c_int32()
    sdpx_gpu.check(L.sdpx_gpu_mp_ldl_dev(nl, n, work.data_ptr(), n, None, 0, None, None,
                                         dinv.data_ptr(), 0, ctypes.byref(ok)))
    assert ok.value == 1 and sdpx_gpu.mp_ldl_stats()["declined"] == 0
    d = mp.to_float64(dinv.cpu().numpy().view(mp.dtype(nl)))
    assert np.all(d > 0)
    assert abs(-np.log2(d).sum() - e.sum()) < 1e-6, (-np.log2(d).sum(), e.sum())
    again = da.clone()
    dinv2 = torch.empty_like(dinv)
    sdpx_gpu.check(L.sdpx_gpu_mp_ldl_dev(nl, n, again.data_ptr(), n, None, 0, None, None,
                                         dinv2.data_ptr(), 0, ctypes.byref(ok)))
    assert ok.value == 1
    assert torch.equal(again, work) and torch.equal(dinv2, dinv)
    del again
    rng = np.random.default_rng(11)
    bmat = rng.standard_normal((n, nrhs))
    dx = dev(mp.from_float64(bmat.T.reshape(-1), nl))
    sdpx_gpu.check(L.sdpx_gpu_mp_ldl_solve_dev(nl, n, work.data_ptr(), n, dinv.data_ptr(), nrhs,
                                               dx.data_ptr(), n))
    del work

    def widen(t, count):  # records of 8 limbs -> 16 limbs, same values
        r = t.view(count, vb)
        return torch.cat([torch.zeros(count, 64, dtype=torch.uint8, device="cuda"), r], dim=1).reshape(-1)

    a16, x16 = widen(da, n * n), widen(dx, n * nrhs)
    del da
    ax16 = torch.empty(n * nrhs * mp.dtype(16).itemsize, dtype=torch.uint8, device="cuda")
    assert sdpx_gpu.rns_gemm_dev("N", "N", n, nrhs, n, a16.data_ptr(), n, x16.data_ptr(), n, 16, False,
                                 ax16.data_ptr(), n * nrhs) == sdpx_gpu.OK
    resid = mp.sub_exact_log2(mp.from_float64(bmat.T.reshape(-1), 16),
                              ax16.cpu().numpy().view(mp.dtype(16)))
    print("residual 2^%.1f" % resid, flush=True)
    assert resid - np.log2(np.abs(bmat).max()) < -400
