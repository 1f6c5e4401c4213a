"""A sampled problem read from an SDPB pmp2sdp JSON directory
(sdpx_gpu.sdpb_json) through the GPU Schur build + arrow factor, bitwise
against the C oracle on the same read records."""
import numpy as np
import pytest

import sdpx_gpu
from helpers import oracle_sampled_arrow
from sdpx_gpu import mp, sdpb_json, workloads
from test_sdpb_json import scalar_directory

pytestmark = pytest.mark.gpu


def test_read_directory_through_the_gpu_factor(gpu, tmp_path):
    nl = 4
    src = workloads.sampled_problem(nl, nblocks=6, hmin=10, hmax=30, t=300, seed=11, lift=False)
    scalar_directory(tmp_path, src)
    parsed = sdpb_json.read_sdpb_sampled(str(tmp_path), nl)
    prob = sdpb_json.arrow_problem(parsed, nl, rinv=[b["rinv"] for b in src["blocks"]],
                                   weights=[b["w"] for b in src["blocks"]], border=src["border"])
    print("sdpb directory", len(prob["blocks"]), prob["t"], flush=True)
    ok, border, bdinv, outs = oracle_sampled_arrow(prob, nl, threads=16)
    okg, bg, dg, fg = sdpx_gpu.sampled_arrow(prob, nl, want_factors=True)
    assert ok and okg and sdpx_gpu.sampled_stats()["declined"] == 0
    for (l, d, y), (lg, dgg, yg) in zip(outs, fg):
        assert np.all(mp.same_bits(lg, l)) and np.all(mp.same_bits(dgg, d)) and np.all(mp.same_bits(yg, y))
    t = prob["t"]
    lower = np.tril_indices(t)
    idx = lower[0] + lower[1] * t
    assert np.all(mp.same_bits(bg[idx], border[idx])) and np.all(mp.same_bits(dg, bdinv))
