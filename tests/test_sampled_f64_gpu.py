"""Float64 sampled Schur build + arrow factor on the GPU.

Float64 parity is by accuracy, not bits: the reference's Float64 path is
LAPACK/BLAS, which fixes no summation order.  The Schur complement
S = C - sum_i B_i^T H_i^-1 B_i cancels heavily and the leaf matrices H_i reach
condition numbers near 1e8, so two correct binary64 evaluations differ far
above 1e-12.  The yardstick is an extended-precision (x87 80-bit,
64-bit significand) evaluation of the same formulas: the GPU must be within
1e-12 of the largest entry of it, or at most as far from it as numpy's own
LAPACK-based binary64 evaluation (times 4)."""
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


def extended(problem):
    """The same S in numpy longdouble (own Cholesky and forward solve)."""
    S = problem["border_f64"].astype(np.longdouble)
    for b in problem["blocks"]:
        V = b["rinv_f64"].astype(np.longdouble) @ b["q_f64"].astype(np.longdouble)
        G = V.T @ V
        w = b["w_f64"].astype(np.longdouble)
        H = np.outer(w, w) * G * G
        K = H.shape[0]
        L = np.zeros_like(H)
        for j in range(K):
            L[j, j] = np.sqrt(H[j, j] - np.sum(L[j, :j] ** 2))
            L[j + 1:, j] = (H[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
        B = b["b_f64"].astype(np.longdouble)
        Y = np.zeros_like(B)
        for i in range(K):
            Y[i] = (B[i] - L[i, :i] @ Y[:i]) / L[i, i]
        S -= Y.T @ Y
    return S


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
    Sx = extended(prob)
    iu = np.triu_indices(t)
    scale = float(np.abs(Sx).max())
    err_gpu = float(np.abs(S[iu] - Sx[iu]).max())
    err_lapack = float(np.abs(Sref[iu] - Sx[iu]).max())
    print(f"S: gpu err {err_gpu / scale:.3e}, numpy err {err_lapack / scale:.3e} (relative to max)")
    assert err_gpu <= max(1e-12 * scale, 4 * err_lapack)
    # the border factor: backward error of -S = F F^T at binary64 level
    Sg = np.triu(S) + np.triu(S, 1).T
    assert np.abs(F @ F.T + Sg).max() <= 1e-12 * np.abs(Sg).max()
