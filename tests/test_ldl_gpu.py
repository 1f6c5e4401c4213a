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
    assert st["gpu_products"] >= 1 and st["regularized"] == count
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
    assert sdpx_gpu.mp_ldl_stats()["gpu_products"] >= 2
    assert ok and okg
    bad = ~mp.same_bits(Lg, L)
    assert not bad.any(), f"{bad.sum()} factor entries differ"
    assert np.all(mp.same_bits(dg, dinv))
