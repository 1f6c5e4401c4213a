"""torch.autograd.Function over the GPU IPM + adjoint (test-only binding)."""
import numpy as np
import torch

import sdpx_gpu
from sdpx_gpu import ipm


class SdpSolve(torch.autograd.Function):
    """y(b, svec(C)) at the central-path point X S = 1e-6 I of the dense SDP;
    backward runs the transposed-KKT adjoint with the GPU factor
    (sdpx_gpu_adjoint_psd_f64), the exact derivative of that point."""

    @staticmethod
    def forward(ctx, b, c_svec, prob):
        p = dict(prob)
        p["b"] = b.detach().cpu().numpy().astype(np.float64)
        p["C"] = ipm.smat(c_svec.detach().cpu().numpy().astype(np.float64), prob["n"])
        X, y, S, info = ipm.solve(p, backend="gpu", target_mu=1e-6, max_iter=200)
        M = sdpx_gpu.schur_psd_f64(info["rinv"], p["a_svec"])
        f = sdpx_gpu.DenseFactor(p["m"])
        assert f.refactor(np.triu(M) + np.triu(M, 1).T)
        ctx.saved = (f, info["rinv"], p["a_svec"])
        return torch.tensor(y, dtype=torch.float64)

    @staticmethod
    def backward(ctx, ybar):
        f, rinv, a_svec = ctx.saved
        bg, cg = sdpx_gpu.adjoint_psd_f64(f, rinv, a_svec,
                                          ybar.detach().cpu().numpy().reshape(-1, 1))
        return (torch.tensor(bg[:, 0]), torch.tensor(cg[:, 0]), None)


class SdpSolveFull(torch.autograd.Function):
    """(y*, svec(X*), <C, X*>) of (b, svec(C), A_svec) at the central-path
    point X S = 1e-6 I; backward runs sdpx_gpu_adjoint_psd_full_f64 (one
    solve with the GPU factor) for the cotangents of all three outputs and
    returns the gradients for b, svec(C) and every svec(A_i)."""

    @staticmethod
    def forward(ctx, b, c_svec, a_svec, prob):
        p = dict(prob)
        n = prob["n"]
        p["b"] = b.detach().cpu().numpy().astype(np.float64)
        p["C"] = ipm.smat(c_svec.detach().cpu().numpy().astype(np.float64), n)
        p["a_svec"] = np.asfortranarray(a_svec.detach().cpu().numpy().astype(np.float64))
        p["A"] = [ipm.smat(p["a_svec"][:, i], n) for i in range(p["a_svec"].shape[1])]
        X, y, S, info = ipm.solve(p, backend="gpu", target_mu=1e-6, max_iter=300)
        assert info["status"] == "centred", info["status"]
        M = sdpx_gpu.schur_psd_f64(info["rinv"], p["a_svec"])
        f = sdpx_gpu.DenseFactor(p["m"])
        assert f.refactor(np.triu(M) + np.triu(M, 1).T)
        ctx.saved = (f, info["rinv"], p["a_svec"], y, ipm.svec(X), ipm.svec(p["C"]))
        return (torch.tensor(y), torch.tensor(ipm.svec(X)),
                torch.tensor(float(np.sum(p["C"] * X)), dtype=torch.float64))

    @staticmethod
    def backward(ctx, ybar, xbar, objbar):
        f, rinv, a_svec, y, x_svec, c_svec = ctx.saved
        bg, cg, ag = sdpx_gpu.adjoint_psd_full_f64(
            f, rinv, a_svec, y, x_svec, c_svec, ybar=ybar.detach().cpu().numpy().reshape(-1, 1),
            xbar=xbar.detach().cpu().numpy().reshape(-1, 1),
            obj=np.array([float(objbar)]), want_a=True)
        return (torch.tensor(bg[:, 0]), torch.tensor(cg[:, 0]), torch.tensor(ag[:, :, 0]), None)
