"""Float64 sampled Schur build + arrow factor on the GPU against a numpy
float64 evaluation of the same formulas (Float64 parity is by tolerance,
1e-12 relative to the largest entry)."""
import numpy as np
import pytest

import sdpx_gpu
from sdpx_gpu import workloads

pytestmark = pytest.mark.gpu


def reference(problem):
    t = problem["t"]
    S = problem["border_f64"].copy()
    for b in problem["blocks"]:
        V = b["rinv_f64"] @ b["q_f64"]
        G = V.T @ V
        H = np.outer(b["w_f64"], b["w_f64"]) * G * G
        L = np.linalg.cholesky(H)
        Y = np.linalg.solve(L, b["b_f64"])
        S -= Y.T @ Y
    return S, np.linalg.cholesky(-S)


@pytest.mark.parametrize("blocks,t", [(4, 150), (40, 2000)])
def test_f64_arrow_matches_numpy(gpu, blocks, t):
    # well-conditioned leaves (square bases, weights in [0.5, 1]): the
    # bootstrap weights of the 256-bit workload make binary64 leaves singular
    prob = workloads.sampled_problem(1, nblocks=blocks, t=t, seed=5, lift=False,
                                     weight_range=(0.5, 1.0), samples=lambda h: h // 2)
    dev = sdpx_gpu.SampledF64(prob)
    info, times = dev.run()
    assert info == 0, info
    S = dev.s.cpu().numpy().reshape(t, t, order="F")
    F = np.tril(dev.factor.cpu().numpy().reshape(t, t, order="F"))
    Sref, Fref = reference(prob)
    iu = np.triu_indices(t)
    assert np.abs(S[iu] - Sref[iu]).max() <= 1e-12 * np.abs(Sref).max()
    assert np.abs(F @ F.T + Sref).max() <= 1e-12 * np.abs(Sref).max()
