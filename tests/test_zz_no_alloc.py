"""The allocation-free exact product (sdpx_gpu_rns_gemm_dev_noalloc) gives
the bits of sdpx_gpu_rns_gemm_dev on device buffers, on two calls in a
row (the second smaller, over the same fixed scratch), and refuses a
product whose temporaries exceed the scratch with a capacity error,
leaving the output untouched."""
import ctypes

import numpy as np
import pytest

import sdpx_gpu
from sdpx_gpu import mp

pytestmark = pytest.mark.gpu


def _dev(rec):
    import torch

    return torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8).reshape(-1).copy()).to("cuda")


def _declare():
    L = sdpx_gpu.lib()
    u8, u32, u64 = ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64
    vp = ctypes.c_void_p
    for name in ("sdpx_gpu_rns_gemm_dev_noalloc", "sdpx_gpu_rns_gemm_dev"):
        getattr(L, name).argtypes = [u32, u8, u8, u64, u64, u64, vp, u64, vp, u64, u32, vp, u64]
        getattr(L, name).restype = ctypes.c_int32
    return L


def _run(L, name, nl, ta, tb, m, n, k, da, lda, db, ldb, upper, out_len):
    import torch

    vb = mp.dtype(nl).itemsize
    out = torch.zeros(out_len * vb, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    code = getattr(L, name)(nl, ord(ta), ord(tb), m, n, k, da.data_ptr(), lda, db.data_ptr(), ldb,
                            1 if upper else 0, out.data_ptr(), out_len)
    return code, out.cpu().numpy().view(mp.dtype(nl))


@pytest.mark.parametrize("nl", [4, 8])
def test_noalloc_matches_workspace_entry_twice(gpu, nl):
    L = _declare()
    rng = np.random.default_rng(nl)
    # the second call smaller than the first, on the same scratch
    cases = ([("N", "T", 300, 280, 257, False), ("T", "N", 120, 120, 90, True)] if nl == 4 else
             [("N", "T", 200, 190, 130, False), ("T", "N", 100, 100, 70, True)])
    for ta, tb, m, n, k, upper in cases:
        print("noalloc", nl, ta, tb, m, n, k, upper, flush=True)
        ar, br = (m if ta == "N" else k), (k if tb == "N" else n)
        a = mp.random_full(ar * (k if ta == "N" else m), nl, 40, rng)
        b = mp.random_full(br * (n if tb == "N" else k), nl, 40, rng)
        da, db = _dev(a), _dev(b)
        out_len = n * (n + 1) // 2 if upper else m * n
        c1, ref = _run(L, "sdpx_gpu_rns_gemm_dev", nl, ta, tb, m, n, k, da, ar, db, br, upper, out_len)
        c2, got = _run(L, "sdpx_gpu_rns_gemm_dev_noalloc", nl, ta, tb, m, n, k, da, ar, db, br, upper,
                       out_len)
        assert c1 == sdpx_gpu.OK and c2 == sdpx_gpu.OK, sdpx_gpu.last_error()
        assert np.all(mp.same_bits(got, ref))


def test_noalloc_rejects_products_beyond_the_scratch(gpu):
    import torch

    L = _declare()
    nl, m, n, k = 4, 1024, 1024, 1024
    rng = np.random.default_rng(3)
    a = mp.random_full(m * k, nl, 20, rng)
    da = _dev(a)
    vb = mp.dtype(nl).itemsize
    out = torch.full((m * n * vb,), 7, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    code = L.sdpx_gpu_rns_gemm_dev_noalloc(nl, ord("N"), ord("T"), m, n, k, da.data_ptr(), m,
                                           da.data_ptr(), m, 0, out.data_ptr(), m * n)
    assert code == sdpx_gpu.CAPACITY
    assert "allocation-free scratch" in sdpx_gpu.last_error()
    assert bool((out == 7).all())
