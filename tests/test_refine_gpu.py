"""Iterative refinement of the sampled arrow KKT solve on the GPU
(sdpx_gpu_sampled_kkt_solve_dev) and the sampled Schur blocks it reads
(sdpx_gpu_sampled_schur_dev), bitwise against the C oracle: the blocks,
the refined solutions, their residual norms, the columns' finiteness and
the number of correction solves, under several stopping settings, on a
reduced instance and on the full 40-block, t = 2000 instance at 256
bits."""
from fractions import Fraction

import numpy as np
import pytest

import pins
import sdpx_gpu
from helpers import oracle_sampled_arrow, oracle_sampled_refine, oracle_sampled_schur
from sdpx_gpu import mp, workloads

pytestmark = pytest.mark.gpu

NL = 4


def run_both(prob, nrhs, settings, seed=3):
    t = prob["t"]
    n = sum(b["k"] for b in prob["blocks"]) + t
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, NL, threads=16)
    h_o = oracle_sampled_schur(prob, NL, threads=16)
    run = sdpx_gpu.SampledDev(prob, NL)
    assert ok and run.run()
    h_g = run.schur()
    assert np.all(mp.same_bits(h_g, h_o)), "sampled Schur blocks differ"
    rng = np.random.default_rng(seed)
    rhs = mp.from_float64(rng.standard_normal(n * nrhs), NL)
    for name, (rt, at, sr, it) in settings.items():
        print("refine", len(prob["blocks"]), t, name, flush=True)
        xg, okg, ng, cg = run.solve_refined(rhs, nrhs, rt, at, sr, it)
        xo, oko, no, co = oracle_sampled_refine(prob, h_o, outs, border, bdinv, rhs, nrhs, rt, at, sr, it,
                                                threads=16)
        print("refine", name, "corrections", cg, co, flush=True)
        assert okg == oko and cg == co, name
        assert np.all(mp.same_bits(ng, no)), name
        bad = ~mp.same_bits(xg, xo)
        assert not bad.any(), f"{name}: {bad.sum()} of {xg.size} entries differ"


TINY = pins.enc([Fraction(1, 2 ** 400)], NL)
HUGE = pins.enc([Fraction(2) ** 1000], NL)


def test_refined_solve_reduced_instance(gpu):
    prob = workloads.sampled_problem(NL, nblocks=6, hmin=10, hmax=40, t=300, seed=17)
    run_both(prob, 3, {"to the stop ratio": (TINY, TINY, 5.0, 10), "no iterations": (TINY, TINY, 5.0, 0),
                       "met at once": (1.0, 1.0, 5.0, 10), "one step": (TINY, TINY, HUGE, 10),
                       "loose": (1e-60, 1e-60, 5.0, 10)})


def test_refined_solve_full_instance(gpu):
    run_both(workloads.sampled_problem(NL, seed=7), 2, {"to the stop ratio": (TINY, TINY, 5.0, 10)})
