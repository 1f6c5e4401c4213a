"""Sampled Schur build + arrow factor with GPU exact products, bitwise
against the C oracle: a reduced instance and the full 40-block,
t = 2000 instance at 256 bits."""
import numpy as np
import pytest

import sdpx_gpu
from helpers import oracle_sampled_arrow
from sdpx_gpu import mp, workloads

pytestmark = pytest.mark.gpu


def compare(prob, nl, threads=16):
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl, threads=threads)
    okg, bg, dg, fg = sdpx_gpu.sampled_arrow(prob, nl, want_factors=True)
    st = sdpx_gpu.sampled_stats()
    print("sampled", nl, len(prob["blocks"]), prob["t"], st, flush=True)
    assert ok and okg
    assert st["gpu_products"] > 0
    for (l, d, y), (lg, dgg, yg) in zip(outs, fg):
        assert np.all(mp.same_bits(lg, l))
        assert np.all(mp.same_bits(dgg, d))
        assert np.all(mp.same_bits(yg, y))
    t = prob["t"]
    lower = np.tril_indices(t)
    idx = lower[0] + lower[1] * t
    bad = ~mp.same_bits(bg[idx], border[idx])
    assert not bad.any(), f"{bad.sum()} border entries differ"
    assert np.all(mp.same_bits(dg, bdinv))


def test_reduced_instance(gpu):
    compare(workloads.sampled_problem(4, nblocks=6, hmin=10, hmax=30, t=400, seed=11), 4)


def test_full_instance_256(gpu):
    compare(workloads.sampled_problem(4, seed=7), 4)
