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
    assert st["gpu_products"] > 0 and st["declined"] == 0, "device pipeline declined"
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


@pytest.mark.parametrize("group", [4, 8, 16, 32])
def test_forward_rows_group_sizes(gpu, group):
    """Y = L^-1 B of the leaves with 4, 8, 16 or 32 lanes per exact dot
    (tuning switch fwd_group): the same bits."""
    sdpx_gpu.set_tuning("fwd_group", group)
    try:
        compare(workloads.sampled_problem(4, nblocks=6, hmin=10, hmax=30, t=400, seed=11), 4)
    finally:
        sdpx_gpu.set_tuning("fwd_group", 8)


@pytest.mark.parametrize("mode", [0, 2])
def test_leaf_ldl_launch_forms(gpu, mode):
    """The leaves' ldl_dots as column launches (0) or as one launch of row
    workgroups with every leaf's rows interleaved (2): the same bits."""
    sdpx_gpu.set_tuning("ldl_one_launch", mode)
    try:
        compare(workloads.sampled_problem(4, nblocks=6, hmin=10, hmax=30, t=400, seed=11), 4)
    finally:
        sdpx_gpu.set_tuning("ldl_one_launch", 1)


def test_full_instance_256(gpu):
    compare(workloads.sampled_problem(4, seed=7), 4)


def test_device_resident_entry_matches_host_entry(gpu):
    """sdpx_gpu_sampled_arrow_dev on HBM-resident records gives the host
    entry's bits (border factor, its D^-1, and every leaf factor), twice in a
    row on the same buffers."""
    prob = workloads.sampled_problem(4, nblocks=5, hmin=10, hmax=24, t=300, seed=5)
    okh, bh, dh, fh = sdpx_gpu.sampled_arrow(prob, 4, want_factors=True)
    run = sdpx_gpu.SampledDev(prob, 4)
    t = prob["t"]
    lower = np.tril_indices(t)
    idx = lower[0] + lower[1] * t
    dt = mp.dtype(4)
    assert sdpx_gpu.sampled_stats()["declined"] == 0
    for rep in range(2):
        print("sampled dev rep", rep, flush=True)
        assert run.run() and okh
        bg, dg = run.result()
        assert np.all(mp.same_bits(bg[idx], bh[idx]))
        assert np.all(mp.same_bits(dg, dh))
        l_all = run.l_all.cpu().numpy().view(dt)
        dinv_all = run.dinv_all.cpu().numpy().view(dt)
        y_all = run.y_all.cpu().numpy().view(dt)
        lo = dlo = ylo = 0
        for b, (l, d, y) in zip(prob["blocks"], fh):
            k = b["k"]
            assert np.all(mp.same_bits(l_all[lo:lo + k * k], l))
            assert np.all(mp.same_bits(dinv_all[dlo:dlo + k], d))
            assert np.all(mp.same_bits(y_all[ylo:ylo + k * t], y))
            lo, dlo, ylo = lo + k * k, dlo + k, ylo + k * t


def test_failed_leaf_index_reported(gpu):
    """A zero weight makes a leaf's Schur block singular: the device
    pipeline reports that leaf's index, as the host panels do."""
    nl = 4
    prob = workloads.sampled_problem(nl, nblocks=5, hmin=8, hmax=14, t=64, seed=19)
    prob["blocks"][2]["w"][3] = mp.zeros(1, nl)[0]
    ok, _, _, _ = sdpx_gpu.sampled_arrow(prob, nl)
    st = sdpx_gpu.sampled_stats()
    assert not ok and st["declined"] == 0 and st["failed_leaf"] == 2
    okh, _, _, _ = sdpx_gpu.sampled_arrow(prob, nl, flags=sdpx_gpu.LDL_HOST_PRODUCTS)
    assert not okh and sdpx_gpu.sampled_stats()["failed_leaf"] == 2


@pytest.mark.parametrize("full", [False, True])
def test_arrow_solve_matches_oracle(gpu, full):
    """ArrowLDLSolver::solve_many on the device factors: bitwise the
    oracle's FMA sweeps and exact coupling dots, on a reduced instance and
    on the full 40-block, t = 2000 instance (4 right-hand sides)."""
    from helpers import oracle_sampled_arrow_solve

    nl = 4
    prob = (workloads.sampled_problem(nl, seed=7) if full else
            workloads.sampled_problem(nl, nblocks=6, hmin=10, hmax=40, t=300, seed=17))
    t, nrhs = prob["t"], 4
    n = sum(b["k"] for b in prob["blocks"]) + t
    print("arrow solve", full, n, flush=True)
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl, threads=16)
    run = sdpx_gpu.SampledDev(prob, nl)
    assert ok and run.run()
    rng = np.random.default_rng(3)
    rhs = mp.from_float64(rng.standard_normal(n * nrhs), nl)
    x = run.solve(rhs, nrhs)
    xo = oracle_sampled_arrow_solve(outs, t, border, bdinv, rhs, nrhs, threads=16)
    bad = ~mp.same_bits(x, xo)
    assert not bad.any(), f"{bad.sum()} of {x.size} solution entries differ"


def test_border_4096_instance(gpu):
    """A 4096-row border (the larger Schur order of the 8192-row target,
    halved to keep the oracle's exact dots to about a minute) with 8
    uneven leaves at 256 bits: bitwise the oracle, no decline."""
    print("sampled t=4096", flush=True)
    compare(workloads.sampled_problem(4, nblocks=8, hmin=10, hmax=60, t=4096, seed=29), 4, threads=16)
