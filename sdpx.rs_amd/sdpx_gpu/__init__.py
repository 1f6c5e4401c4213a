"""Host-side mirror of the SDPX exact-BLAS and dense factor/solve hooks,
bound to the MI355X library ``libsdpx_gpu.so`` through its C ABI
(``include/sdpx_gpu.h``).

The names follow the reference's trait methods (crates/solver/src/algebra/
dense/blas/traits.rs:228-370): ``xgemm_upper_exact`` returns ``False`` on a
decline, exactly like the trait, and leaves the output untouched.  Values
are numpy record arrays in the MpFloat<N> interchange layout (see
``sdpx_gpu.mp``).

The library is required: every entry point raises ``LibraryMissing`` when
``libsdpx_gpu.so`` has not been built, and ``DeviceError`` when the HIP call
fails.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import mp

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libsdpx_gpu.so")

OK, INVALID, DECLINED, BUSY, DEVICE_ERROR, CAPACITY = 0, 1, 2, 3, 5, 6


class LibraryMissing(RuntimeError):
    pass


class DeviceError(RuntimeError):
    pass


class InvalidInput(ValueError):
    pass


_lib = None


def lib():
    """The loaded C library (ctypes.CDLL); raises LibraryMissing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    u8, u32, u64, i32 = ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    L.sdpx_gpu_abi_version.restype = u32
    L.sdpx_gpu_init.argtypes = [i32]
    L.sdpx_gpu_init.restype = i32
    L.sdpx_gpu_stream.restype = vp
    L.sdpx_gpu_last_error.argtypes = [ctypes.c_char_p, u64]
    L.sdpx_gpu_last_error.restype = u64
    L.sdpx_gpu_last_plan.argtypes = [ctypes.POINTER(u32), ctypes.POINTER(u32), dp]
    L.sdpx_gpu_last_plan.restype = i32
    L.sdpx_gpu_set_profiling.argtypes = [i32]
    L.sdpx_gpu_set_profiling.restype = i32
    L.sdpx_gpu_set_tuning.argtypes = [ctypes.c_char_p, ctypes.c_int64]
    L.sdpx_gpu_set_tuning.restype = i32
    L.sdpx_gpu_last_stage_times.argtypes = [dp, u32]
    L.sdpx_gpu_last_stage_times.restype = i32
    L.sdpx_gpu_dense_create.argtypes = [u64, ctypes.POINTER(vp)]
    L.sdpx_gpu_dense_destroy.argtypes = [vp]
    L.sdpx_gpu_dense_set_dynamic_pivot.argtypes = [vp, i32, ctypes.c_double, ctypes.c_double]
    for name in ("sdpx_gpu_dense_refactor", "sdpx_gpu_dense_refactor_dev"):
        getattr(L, name).argtypes = [vp, vp, u64, ctypes.POINTER(i32)]
    for name in ("sdpx_gpu_dense_solve_many", "sdpx_gpu_dense_solve_many_dev"):
        getattr(L, name).argtypes = [vp, vp, u64, u64]
    L.sdpx_gpu_dense_factor_dev_ptr.argtypes = [vp, ctypes.POINTER(vp)]
    L.sdpx_gpu_dense_get_factor.argtypes = [vp, vp, u64]
    L.sdpx_gpu_dense_get_factor.restype = i32
    for name in ("sdpx_gpu_schur_psd_f64", "sdpx_gpu_schur_psd_f64_dev"):
        getattr(L, name).argtypes = [u64, u64, vp, vp, vp, u64]
    for name in ("sdpx_gpu_dense_create", "sdpx_gpu_dense_destroy", "sdpx_gpu_dense_set_dynamic_pivot",
                 "sdpx_gpu_dense_refactor", "sdpx_gpu_dense_refactor_dev", "sdpx_gpu_dense_solve_many",
                 "sdpx_gpu_dense_solve_many_dev", "sdpx_gpu_dense_factor_dev_ptr",
                 "sdpx_gpu_schur_psd_f64", "sdpx_gpu_schur_psd_f64_dev"):
        getattr(L, name).restype = i32
    L.sdpx_gpu_mp_ldl.argtypes = [u32, u64, vp, u64, vp, u64, vp, vp, vp, u32,
                                  ctypes.POINTER(i32)]
    L.sdpx_gpu_mp_ldl.restype = i32
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.sdpx_gpu_mp_ldl_stats.argtypes = [i64p, i64p, i64p, dp, i64p, i64p]
    L.sdpx_gpu_mp_ldl_stats.restype = i32
    L.sdpx_gpu_sampled_arrow.argtypes = [u32, u64, vp, u64, vp, vp, u64, vp, vp, u32,
                                         ctypes.POINTER(i32)]
    L.sdpx_gpu_sampled_arrow.restype = i32
    L.sdpx_gpu_sampled_stats.argtypes = [dp] + [ctypes.POINTER(ctypes.c_int64)] * 5
    L.sdpx_gpu_panel_phase_times.argtypes = [vp, vp, i32]
    L.sdpx_gpu_panel_phase_times.restype = i32
    L.sdpx_gpu_sampled_arrow_solve_dev.argtypes = [u32, u64, vp, u64, vp, vp, vp, vp, vp, u64, vp, u64]
    L.sdpx_gpu_sampled_arrow_solve_dev.restype = i32
    L.sdpx_gpu_psd_nt_scaling_f64_dev.argtypes = [u64, vp, vp, vp, vp, vp, vp, vp, vp]
    L.sdpx_gpu_psd_nt_scaling_f64_dev.restype = i32
    L.sdpx_gpu_sampled_schur_dev.argtypes = [u32, u64, vp, vp]
    L.sdpx_gpu_sampled_schur_dev.restype = i32
    L.sdpx_gpu_sampled_kkt_solve_dev.argtypes = [u32, u64, vp, u64, vp, vp, vp, vp, vp, vp, vp, vp, u64, vp, u64,
                                                 vp, u64, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.sdpx_gpu_sampled_kkt_solve_dev.restype = i32
    L.sdpx_gpu_sampled_stats.restype = i32
    L.sdpx_gpu_sampled_arrow_f64_dev.argtypes = [u64, vp, u64, vp, vp, vp, ctypes.c_double,
                                                 ctypes.c_double, ctypes.POINTER(i32), dp]
    L.sdpx_gpu_sampled_arrow_f64_dev.restype = i32
    L.sdpx_gpu_rns_totals.argtypes = [i32, dp, dp, dp, ctypes.POINTER(ctypes.c_int64)]
    L.sdpx_gpu_rns_totals.restype = i32
    for name in ("sdpx_gpu_adjoint_psd_f64", "sdpx_gpu_adjoint_psd_f64_dev"):
        getattr(L, name).argtypes = [vp, u64, vp, vp, u64, vp, vp, vp]
        getattr(L, name).restype = i32
    for name in ("sdpx_gpu_adjoint_psd_full_f64", "sdpx_gpu_adjoint_psd_full_f64_dev"):
        getattr(L, name).argtypes = [vp, u64, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, vp, vp]
        getattr(L, name).restype = i32
    L.sdpx_gpu_condensed_assemble_f64_dev.argtypes = [u64, vp, vp, vp, u64, vp, u64, vp, u64, vp, vp, u64]
    L.sdpx_gpu_condensed_assemble_f64_dev.restype = i32
    for name in ("sdpx_gpu_condensed_rhs_f64_dev", "sdpx_gpu_condensed_recover_f64_dev"):
        getattr(L, name).argtypes = [u64, u64, vp, u64, vp, u64, vp, vp, vp, vp]
        getattr(L, name).restype = i32
    L.sdpx_gpu_rns_gemm_blocks_upper.argtypes = [u32, u64, u64, vp, vp]
    L.sdpx_gpu_rns_gemm_blocks_upper.restype = i32
    L.sdpx_gpu_mp_ldl_dev.argtypes = [u32, u64, vp, u64, vp, u64, vp, vp, vp, u32,
                                      ctypes.POINTER(i32)]
    L.sdpx_gpu_mp_ldl_dev.restype = i32
    for name in ("sdpx_gpu_mp_ldl_solve", "sdpx_gpu_mp_ldl_solve_dev"):
        getattr(L, name).argtypes = [u32, u64, vp, u64, vp, u64, vp, u64]
        getattr(L, name).restype = i32
    L.sdpx_gpu_sampled_arrow_dev.argtypes = [u32, u64, vp, u64, vp, vp, vp, vp, vp, u64, vp,
                                             ctypes.POINTER(i32)]
    L.sdpx_gpu_sampled_arrow_dev.restype = i32
    L.sdpx_gpu_mp_potrf.argtypes = [u32, u8, u64, vp, u64, ctypes.POINTER(i32)]
    L.sdpx_gpu_mp_potrf.restype = i32
    L.sdpx_gpu_mp_potrs.argtypes = [u32, u8, u64, u64, vp, u64, vp, u64, ctypes.POINTER(i32)]
    L.sdpx_gpu_mp_potrs.restype = i32
    L.sdpx_gpu_mp_trsm_lower.argtypes = [u32, u64, vp, vp, u64, u32]
    L.sdpx_gpu_mp_trsm_lower.restype = i32
    gemm_args = [u32, u8, u8, u64, u64, u64, vp, u64, vp, u64, u32, vp, u64]
    for name in ("sdpx_gpu_rns_gemm", "sdpx_gpu_rns_gemm_dev"):
        f = getattr(L, name)
        f.argtypes = gemm_args
        f.restype = i32
    for name in ("sdpx_gpu_xgemm", "sdpx_gpu_xgemm_dev"):
        getattr(L, name).argtypes = [u32, u8, u8, u64, u64, u64, vp, vp, u64, vp, u64, vp, vp, u64]
        getattr(L, name).restype = i32
    for name in ("sdpx_gpu_xsyrk", "sdpx_gpu_xsyrk_dev"):
        getattr(L, name).argtypes = [u32, u8, u8, u64, u64, vp, vp, u64, vp, vp, u64, u64]
        getattr(L, name).restype = i32
    L.sdpx_gpu_diag_congruence_upper_exact.argtypes = [u32, u64, u64, vp, vp, vp]
    L.sdpx_gpu_diag_congruence_upper_exact_dev.argtypes = [u32, u64, u64, vp, vp, vp, u32]
    for name in ("sdpx_gpu_congruence_exact", "sdpx_gpu_congruence_exact_dev"):
        getattr(L, name).argtypes = [u32, u8, u64, u64, vp, u64, vp, u64, vp, u32]
    for name in ("sdpx_gpu_svec_quadratic_exact", "sdpx_gpu_svec_quadratic_exact_dev"):
        getattr(L, name).argtypes = [u32, u64, u64, vp, u32, vp, vp, vp]
    for name in ("sdpx_gpu_symmetric_bilinear_exact", "sdpx_gpu_symmetric_bilinear_exact_dev"):
        getattr(L, name).argtypes = [u32, u64, u64, vp, vp, u64, vp, vp]
    for name in ("sdpx_gpu_diag_congruence_upper_exact", "sdpx_gpu_diag_congruence_upper_exact_dev",
                 "sdpx_gpu_congruence_exact", "sdpx_gpu_congruence_exact_dev", "sdpx_gpu_svec_quadratic_exact",
                 "sdpx_gpu_svec_quadratic_exact_dev", "sdpx_gpu_symmetric_bilinear_exact",
                 "sdpx_gpu_symmetric_bilinear_exact_dev"):
        getattr(L, name).restype = i32


def last_error():
    L = lib()
    need = L.sdpx_gpu_last_error(None, 0)
    buf = ctypes.create_string_buffer(int(need))
    L.sdpx_gpu_last_error(buf, need)
    return buf.value.decode(errors="replace")


def check(code, allow_decline=False):
    if code == OK or (allow_decline and code == DECLINED):
        return code
    msg = last_error()
    if code == INVALID:
        raise InvalidInput(msg)
    raise DeviceError(f"sdpx_gpu error {code}: {msg}")


def last_plan():
    """(moduli, planes, log2 M) of this thread's most recent exact product."""
    L = lib()
    a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
    L.sdpx_gpu_last_plan(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def set_profiling(on):
    lib().sdpx_gpu_set_profiling(1 if on else 0)


def set_tuning(name, value):
    """A named tuning switch of the panel kernels (sdpx_gpu_set_tuning)."""
    check(lib().sdpx_gpu_set_tuning(name.encode(), int(value)))


def last_stage_times():
    """Stage times (ms) of this thread's last exact product: scan, encode,
    product, crt (needs set_profiling(True))."""
    buf = (ctypes.c_double * 4)()
    lib().sdpx_gpu_last_stage_times(buf, 4)
    return list(buf)


def rns_totals(reset=False):
    """Totals over this thread's profiled exact products (see
    sdpx_gpu_rns_totals)."""
    ms = (ctypes.c_double * 4)()
    ops, macs = ctypes.c_double(), ctypes.c_double()
    calls = ctypes.c_int64()
    lib().sdpx_gpu_rns_totals(1 if reset else 0, ms, ctypes.byref(ops), ctypes.byref(macs),
                              ctypes.byref(calls))
    return {"stage_ms": list(ms), "int8_ops": ops.value, "mp_macs": macs.value,
            "calls": calls.value}


def _ptr(arr):
    return arr.ctypes.data_as(ctypes.c_void_p) if arr is not None and arr.size else None


def rns_gemm(ta, tb, m, n, k, a, lda, b, ldb, upper_only=False, out=None, out_len=None):
    """rns_blas::gemm (rns_blas.rs:2134-2417) on host record arrays.

    Returns ``(ok, out)``: ``ok`` is False on a decline (out untouched)."""
    nl = mp.limbs_of(a if a is not None and a.size else b)
    if out_len is None:
        out_len = n * (n + 1) // 2 if (upper_only and m == n) else m * n
    if out is None:
        out = mp.zeros(out_len, nl)
    code = lib().sdpx_gpu_rns_gemm(nl, ord(ta), ord(tb), m, n, k, _ptr(a), lda, _ptr(b), ldb,
                                   1 if upper_only else 0, _ptr(out), out_len)
    check(code, allow_decline=True)
    return code == OK, out


def xgemm_upper_exact(ta, tb, m, n, k, a, lda, b, ldb, c):
    """XgemmScalar::xgemm_upper_exact (traits.rs:240-249): upper triangle of
    op(a) op(b) into c (column-packed when square with n(n+1)/2 entries,
    else ldc = m).  Returns False on a decline with c untouched."""
    ok, _ = rns_gemm(ta, tb, m, n, k, a, lda, b, ldb, True, c, c.size)
    return ok


def rns_gemm_dev(ta, tb, m, n, k, a_ptr, lda, b_ptr, ldb, limbs, upper_only, out_ptr, out_len):
    """Device-pointer form (pointers as ints); returns the status code."""
    code = lib().sdpx_gpu_rns_gemm_dev(limbs, ord(ta), ord(tb), m, n, k, a_ptr, lda, b_ptr, ldb,
                                       1 if upper_only else 0, out_ptr, out_len)
    return check(code, allow_decline=True)


class DenseFactor:
    """Float64 dense factor/solve on the GPU, mirroring DirectLDLSolver's
    refactor / solve_many (crates/solver/src/solver/kkt/direct/mod.rs:11-62)
    for DenseBlockSolver (kkt/ldl/dense_block.rs:784-851)."""

    def __init__(self, n):
        self.n = n
        h = ctypes.c_void_p()
        check(lib().sdpx_gpu_dense_create(n, ctypes.byref(h)))
        self._h = h
        self.info = 0

    def set_dynamic_pivot(self, eps, delta, enable=True):
        check(lib().sdpx_gpu_dense_set_dynamic_pivot(self._h, 1 if enable else 0, eps, delta))

    def refactor(self, a):
        """Factor the symmetric matrix whose upper triangle ``a`` holds
        (column-major n x n).  Returns False on a failed pivot (info > 0)."""
        a = np.asfortranarray(a, dtype=np.float64)
        info = ctypes.c_int32()
        check(lib().sdpx_gpu_dense_refactor(self._h, a.ctypes.data_as(ctypes.c_void_p), self.n,
                                            ctypes.byref(info)))
        self.info = info.value
        return self.info == 0

    def refactor_dev(self, ptr, lda):
        info = ctypes.c_int32()
        check(lib().sdpx_gpu_dense_refactor_dev(self._h, ptr, lda, ctypes.byref(info)))
        self.info = info.value
        return self.info == 0

    def solve_many(self, b):
        """A^-1 b for an n x nrhs panel (returns a new array)."""
        b = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        check(lib().sdpx_gpu_dense_solve_many(self._h, b.ctypes.data_as(ctypes.c_void_p), nrhs,
                                              self.n))
        return b

    def solve_many_dev(self, ptr, nrhs, ldb):
        check(lib().sdpx_gpu_dense_solve_many_dev(self._h, ptr, nrhs, ldb))

    def factor(self):
        """The lower factor L (A = L L^T) copied to the host."""
        out = np.zeros((self.n, self.n), order="F")
        check(lib().sdpx_gpu_dense_get_factor(self._h, out.ctypes.data_as(ctypes.c_void_p), self.n))
        return np.tril(out)

    def factor_ptr(self):
        p = ctypes.c_void_p()
        check(lib().sdpx_gpu_dense_factor_dev_ptr(self._h, ctypes.byref(p)))
        return p.value

    def close(self):
        if self._h:
            lib().sdpx_gpu_dense_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def schur_psd_f64(rinv, a_svec):
    """Upper triangle of S_ab = svec(Rinv A_a Rinv^T) . svec(Rinv A_b Rinv^T)
    (kkt/condensed/psd.rs:1299-1456).  ``a_svec`` is s x m."""
    rinv = np.asfortranarray(rinv, dtype=np.float64)
    a_svec = np.asfortranarray(a_svec, dtype=np.float64)
    n = rinv.shape[0]
    m = a_svec.shape[1]
    S = np.zeros((m, m), order="F")
    check(lib().sdpx_gpu_schur_psd_f64(n, m, rinv.ctypes.data_as(ctypes.c_void_p),
                                       a_svec.ctypes.data_as(ctypes.c_void_p),
                                       S.ctypes.data_as(ctypes.c_void_p), m))
    return S


def svec_f64(mat):
    """The svec of a symmetric matrix: its upper triangle column by column,
    off-diagonals times sqrt(2) (the reference's svec order)."""
    n = mat.shape[0]
    iu = np.triu_indices(n)
    order = np.lexsort((iu[0], iu[1]))
    i, j = iu[0][order], iu[1][order]
    return np.where(i == j, 1.0, np.sqrt(2.0)) * mat[i, j]


class NtScaling:
    """Device buffers for sdpx_gpu_psd_nt_scaling_f64_dev over blocks of the
    given orders; run(s, z) takes the svec'd slacks and duals (concatenated
    device tensors) and leaves R, Rinv, lambda and G in the tensors of the
    same names."""

    def __init__(self, dims):
        import torch

        self.dims = [int(n) for n in dims]
        sv = sum(n * (n + 1) // 2 for n in self.dims)
        mats = sum(n * n for n in self.dims)
        f64 = dict(dtype=torch.float64, device="cuda")
        self.s = torch.zeros(max(1, sv), **f64)
        self.z = torch.zeros(max(1, sv), **f64)
        self.r = torch.zeros(max(1, mats), **f64)
        self.rinv = torch.zeros(max(1, mats), **f64)
        self.g = torch.zeros(max(1, mats), **f64)
        self.lam = torch.zeros(max(1, sum(self.dims)), **f64)
        self._dims = (ctypes.c_uint64 * max(1, len(self.dims)))(*self.dims)
        self.status = (ctypes.c_int32 * max(1, len(self.dims)))()

    def run(self):
        import torch

        torch.cuda.current_stream().synchronize()
        vp = ctypes.c_void_p
        check(lib().sdpx_gpu_psd_nt_scaling_f64_dev(len(self.dims), self._dims, vp(self.s.data_ptr()),
                                                    vp(self.z.data_ptr()), vp(self.r.data_ptr()),
                                                    vp(self.rinv.data_ptr()), vp(self.lam.data_ptr()),
                                                    vp(self.g.data_ptr()), self.status))
        return [int(self.status[i]) for i in range(len(self.dims))]


def psd_nt_scaling_f64(pairs):
    """PSDTriangleCone::update_scaling of several blocks on the GPU: pairs of
    (S, Z) symmetric matrices -> [(R, Rinv, lambda, G, status)] per block."""
    import torch

    dims = [np.asarray(S).shape[0] for S, _ in pairs]
    nt = NtScaling(dims)
    if pairs:
        nt.s.copy_(torch.from_numpy(np.concatenate([svec_f64(np.asarray(S, np.float64)) for S, _ in pairs])))
        nt.z.copy_(torch.from_numpy(np.concatenate([svec_f64(np.asarray(Z, np.float64)) for _, Z in pairs])))
    status = nt.run()
    r, rinv, g, lam = (t.cpu().numpy() for t in (nt.r, nt.rinv, nt.g, nt.lam))
    out, mo, vo = [], 0, 0
    for n, st in zip(dims, status):
        mat = lambda a: a[mo:mo + n * n].reshape(n, n).T.copy()
        out.append((mat(r), mat(rinv), lam[vo:vo + n].copy(), mat(g), st))
        mo += n * n
        vo += n
    return out


LDL_DOTS, LDL_HOST_PRODUCTS = 1, 2


def mp_ldl(a, n, signs=None, reg=None, flags=0):
    """DenseLeaf::factor_leaf (arrow.rs:808-926) on the GPU path: returns
    (ok, l, dinv) with l the unit lower factor in the lower triangle."""
    nl = mp.limbs_of(a)
    l = a.copy()
    dinv = mp.zeros(n, nl)
    sg = None if signs is None else np.ascontiguousarray(signs, np.int8)
    eps = delta = None
    if reg is not None:
        eps, delta = reg
    ok = ctypes.c_int32()
    check(lib().sdpx_gpu_mp_ldl(nl, n, _ptr(l), n, _ptr(sg) if sg is not None else None,
                                0 if sg is None else sg.size, _ptr(eps), _ptr(delta), _ptr(dinv),
                                flags, ctypes.byref(ok)))
    return bool(ok.value), l, dinv


def mp_ldl_stats():
    """Counters of this thread's last mp_ldl: ``declined`` is 1 when the
    device panels declined and the host panels produced the factor,
    ``retried`` 1 when the recursive factor failed and exact dots redid it."""
    a, b, c, e, f = (ctypes.c_int64() for _ in range(5))
    d = ctypes.c_double()
    lib().sdpx_gpu_mp_ldl_stats(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d),
                                ctypes.byref(e), ctypes.byref(f))
    return {"gpu_products": a.value, "host_products": b.value, "regularized": c.value,
            "gpu_mp_macs": d.value, "declined": e.value, "retried": f.value}


class _Sized(ctypes.Structure):
    """An ABI struct whose first member, struct_size, is filled in."""

    def __init__(self, *args, **kw):
        super().__init__(ctypes.sizeof(self), *args, **kw)


class SampledLeaf(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("side", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("rinv", ctypes.c_void_p), ("basis", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("coupling", ctypes.c_void_p)]


class RefineSettings(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("reltol", ctypes.c_void_p), ("abstol", ctypes.c_void_p),
                ("stop_ratio", ctypes.c_void_p), ("max_iter", ctypes.c_uint32)]


class SampledFactor(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("l", ctypes.c_void_p), ("dinv", ctypes.c_void_p), ("y", ctypes.c_void_p)]


def sampled_arrow(problem, nl, border_signs=(-1,), flags=0, want_factors=False):
    """One Schur build + arrow factor of a sampled problem
    (workloads.sampled_problem).  Returns (ok, border, border_dinv, factors)."""
    blocks = problem["blocks"]
    t = problem["t"]
    leaves = (SampledLeaf * len(blocks))()
    for i, b in enumerate(blocks):
        leaves[i] = SampledLeaf(b["h"], b["k"], _ptr(b["rinv"]), _ptr(b["q"]), _ptr(b["w"]),
                                _ptr(b["b"]))
    border = problem["border"].copy()
    bdinv = mp.zeros(t, nl)
    signs = np.ascontiguousarray(border_signs, np.int8)
    facs = None
    outs = []
    if want_factors:
        facs = (SampledFactor * len(blocks))()
        for i, b in enumerate(blocks):
            l, d, y = mp.zeros(b["k"] ** 2, nl), mp.zeros(b["k"], nl), mp.zeros(b["k"] * t, nl)
            outs.append((l, d, y))
            facs[i] = SampledFactor(_ptr(l), _ptr(d), _ptr(y))
    ok = ctypes.c_int32()
    check(lib().sdpx_gpu_sampled_arrow(nl, len(blocks), leaves, t, _ptr(border), _ptr(signs),
                                       signs.size, _ptr(bdinv), facs, flags, ctypes.byref(ok)))
    return bool(ok.value), border, bdinv, outs


_lib_stream = None


def library_stream():
    """The library's HIP stream as a torch stream object: work queued under
    ``with torch.cuda.stream(library_stream())`` is ordered with the
    ``_dev`` entry points, which run on that stream."""
    global _lib_stream
    if _lib_stream is None:
        import torch

        _lib_stream = torch.cuda.ExternalStream(lib().sdpx_gpu_stream(), device="cuda")
    return _lib_stream


class SampledDev:
    """Device-resident multiple-precision form of a sampled problem (torch
    uint8 tensors on cuda holding the interchange records) for
    sdpx_gpu_sampled_arrow_dev.  run() factors a fresh copy of the border;
    the leaf factors land in l_all / dinv_all / y_all in leaf order."""

    def __init__(self, problem, nl, border_signs=(-1,)):
        import torch

        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).to("cuda")

        blocks = problem["blocks"]
        self.nl, self.t = nl, problem["t"]
        self.keep = []
        self.leaves = (SampledLeaf * len(blocks))()
        for i, b in enumerate(blocks):
            arrs = [dev(b[k]) for k in ("rinv", "q", "w", "b")]
            self.keep.append(arrs)
            self.leaves[i] = SampledLeaf(b["h"], b["k"], *[a.data_ptr() for a in arrs])
        vb = mp.dtype(nl).itemsize
        ks = [b["k"] for b in blocks]
        self.ks = ks
        self.border0 = dev(problem["border"])
        self.border = torch.empty_like(self.border0)
        self.bdinv = torch.zeros(self.t * vb, dtype=torch.uint8, device="cuda")
        self.l_all = torch.zeros(max(1, sum(k * k for k in ks)) * vb, dtype=torch.uint8, device="cuda")
        self.dinv_all = torch.zeros(max(1, sum(ks)) * vb, dtype=torch.uint8, device="cuda")
        self.y_all = torch.zeros(max(1, sum(ks) * self.t) * vb, dtype=torch.uint8, device="cuda")
        self.signs = np.ascontiguousarray(border_signs, np.int8)
        torch.cuda.synchronize()

    def run(self):
        """One Schur build + arrow factor; returns ok (False: a pivot failed).
        Raises on a decline (the device limits were exceeded).  The fresh
        border is copied on the library's own stream, so the factor is
        ordered after it."""
        import torch

        with torch.cuda.stream(library_stream()):
            self.border.copy_(self.border0)
        ok = ctypes.c_int32()
        check(lib().sdpx_gpu_sampled_arrow_dev(self.nl, len(self.leaves), self.leaves, self.t,
                                               self.l_all.data_ptr(), self.dinv_all.data_ptr(),
                                               self.y_all.data_ptr(), self.border.data_ptr(),
                                               _ptr(self.signs), self.signs.size,
                                               self.bdinv.data_ptr(), ctypes.byref(ok)))
        return bool(ok.value)

    def solve(self, rhs, nrhs):
        """ArrowLDLSolver::solve_many on the factors of the last run():
        rhs (n x nrhs records, column-major, n = sum K + t: leaf rows in
        leaf order, then the border rows) -> x, same layout."""
        import torch

        nl = self.nl
        n = sum(self.ks) + self.t
        dt = mp.dtype(nl)
        rhs = np.ascontiguousarray(rhs, dtype=dt).reshape(-1)
        assert rhs.size == n * nrhs
        dx = torch.from_numpy(rhs.view(np.uint8).copy()).to("cuda")
        torch.cuda.synchronize()
        ks = (ctypes.c_uint64 * max(1, len(self.ks)))(*self.ks)
        check(lib().sdpx_gpu_sampled_arrow_solve_dev(nl, len(self.ks), ks, self.t, self.l_all.data_ptr(),
                                                     self.dinv_all.data_ptr(), self.y_all.data_ptr(),
                                                     self.border.data_ptr(), self.bdinv.data_ptr(), nrhs,
                                                     dx.data_ptr(), n))
        return dx.cpu().numpy().view(dt).copy()

    def schur(self):
        """The leaves' sampled Schur blocks (sdpx_gpu_sampled_schur_dev):
        H of every leaf (K x K, lower triangle, zeros above) concatenated in
        leaf order, as a device tensor (kept for solve_refined) and host
        records."""
        import torch

        vb = mp.dtype(self.nl).itemsize
        self.h_all = torch.zeros(max(1, sum(k * k for k in self.ks)) * vb, dtype=torch.uint8, device="cuda")
        check(lib().sdpx_gpu_sampled_schur_dev(self.nl, len(self.leaves), self.leaves, self.h_all.data_ptr()))
        return self.h_all.cpu().numpy().view(mp.dtype(self.nl)).copy()

    def solve_refined(self, rhs, nrhs, reltol, abstol, stop_ratio=5.0, max_iter=10):
        """DirectLDLKKTSolver::solve_many with iterative refinement against
        the unshifted KKT (sdpx_gpu_sampled_kkt_solve_dev) on the factors of
        the last run() and the blocks of schur(): rhs (n x nrhs records) ->
        (x, ok per column, residual norms, correction solves).  reltol,
        abstol, stop_ratio: records (or floats, converted exactly)."""
        import torch

        nl = self.nl
        n = sum(self.ks) + self.t
        dt = mp.dtype(nl)

        def rec(v):
            return v if isinstance(v, np.ndarray) else mp.from_float64(np.array([float(v)]), nl)

        rt, at, sr = rec(reltol), rec(abstol), rec(stop_ratio)
        settings = RefineSettings(reltol=_ptr(rt), abstol=_ptr(at), stop_ratio=_ptr(sr), max_iter=max_iter)
        rhs = np.ascontiguousarray(rhs, dtype=dt).reshape(-1)
        assert rhs.size == n * nrhs
        db = torch.from_numpy(rhs.view(np.uint8).copy()).to("cuda")
        dx = torch.zeros_like(db)
        torch.cuda.synchronize()
        ok = (ctypes.c_int32 * max(1, nrhs))()
        norms = mp.zeros(max(1, nrhs), nl)
        count = ctypes.c_int64()
        check(lib().sdpx_gpu_sampled_kkt_solve_dev(
            nl, len(self.leaves), self.leaves, self.t, self.h_all.data_ptr(), self.border0.data_ptr(),
            self.l_all.data_ptr(), self.dinv_all.data_ptr(), self.y_all.data_ptr(), self.border.data_ptr(),
            self.bdinv.data_ptr(), ctypes.byref(settings), nrhs, db.data_ptr(), n, dx.data_ptr(), n, ok,
            _ptr(norms), ctypes.byref(count)))
        return (dx.cpu().numpy().view(dt).copy(), [bool(ok[c]) for c in range(nrhs)], norms[:nrhs],
                count.value)

    def result(self):
        """(border factor, border dinv) as host record arrays."""
        dt = mp.dtype(self.nl)
        return (self.border.cpu().numpy().view(dt).copy(), self.bdinv.cpu().numpy().view(dt).copy())


def panel_phase_times(reset=False):
    """Kernel-category times (ms) of this thread's device pipelines while
    profiling is on: ldl_dots, trsm bases, forward rows, exact products."""
    ms = (ctypes.c_double * 4)()
    n = (ctypes.c_int64 * 4)()
    check(lib().sdpx_gpu_panel_phase_times(ms, n, 1 if reset else 0))
    names = ("ldl_dots", "trsm_base", "forward_rows", "exact_products")
    return {k: ms[i] for i, k in enumerate(names)}, {k: n[i] for i, k in enumerate(names)}


def sampled_stats():
    """Phase times and counters of this thread's last sampled arrow call;
    ``declined`` 1 when the device pipeline declined and host panels ran."""
    secs = (ctypes.c_double * 3)()
    g, h, dc, rt, fl = (ctypes.c_int64() for _ in range(5))
    lib().sdpx_gpu_sampled_stats(secs, ctypes.byref(g), ctypes.byref(h), ctypes.byref(dc),
                                 ctypes.byref(rt), ctypes.byref(fl))
    return {"build_s": secs[0], "leaves_s": secs[1], "border_s": secs[2],
            "gpu_products": g.value, "host_products": h.value, "declined": dc.value,
            "retried": rt.value, "failed_leaf": fl.value}


class SampledLeafF64(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("side", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("rinv", ctypes.c_void_p), ("basis", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("coupling", ctypes.c_void_p)]


class SampledF64:
    """Device-resident float64 form of a sampled problem (torch tensors on
    cuda) for sdpx_gpu_sampled_arrow_f64_dev."""

    def __init__(self, problem):
        import torch

        dev = "cuda"
        self.t = problem["t"]
        self.keep = []
        self.leaves = (SampledLeafF64 * len(problem["blocks"]))()
        for i, b in enumerate(problem["blocks"]):
            arrs = [torch.tensor(np.asfortranarray(b[k]).ravel(order="F"), dtype=torch.float64,
                                 device=dev) for k in ("rinv_f64", "q_f64", "w_f64", "b_f64")]
            self.keep.append(arrs)
            self.leaves[i] = SampledLeafF64(b["h"], b["k"], *[a.data_ptr() for a in arrs])
        self.c = torch.tensor(np.asfortranarray(problem["border_f64"]).ravel(order="F"),
                              dtype=torch.float64, device=dev)
        self.s = torch.zeros_like(self.c)
        self.factor = torch.zeros_like(self.c)
        torch.cuda.synchronize()

    def run(self, dynamic=None):
        """dynamic: (eps, delta) of the dynamic pivot rule, or None."""
        info = ctypes.c_int32()
        times = (ctypes.c_double * 3)()
        eps, delta = dynamic if dynamic else (0.0, 0.0)
        check(lib().sdpx_gpu_sampled_arrow_f64_dev(len(self.leaves), self.leaves, self.t,
                                                   self.c.data_ptr(), self.s.data_ptr(),
                                                   self.factor.data_ptr(), eps, delta,
                                                   ctypes.byref(info), times))
        return info.value, list(times)


def adjoint_psd_f64(factor, rinv, a_svec, ybar):
    """Adjoint of the PSD Schur solve with the GPU factor of M: returns
    (bgrad (m x r), cgrad (s x r)) for cotangents ybar (m x r) of y*."""
    rinv = np.asfortranarray(rinv, dtype=np.float64)
    a_svec = np.asfortranarray(a_svec, dtype=np.float64)
    ybar = np.asfortranarray(np.atleast_2d(ybar.T).T if ybar.ndim == 1 else ybar, dtype=np.float64)
    n = rinv.shape[0]
    m, r = ybar.shape
    s = n * (n + 1) // 2
    bgrad = np.zeros((m, r), order="F")
    cgrad = np.zeros((s, r), order="F")
    vp = ctypes.c_void_p
    check(lib().sdpx_gpu_adjoint_psd_f64(factor._h, n, rinv.ctypes.data_as(vp),
                                         a_svec.ctypes.data_as(vp), r, ybar.ctypes.data_as(vp),
                                         bgrad.ctypes.data_as(vp), cgrad.ctypes.data_as(vp)))
    return bgrad, cgrad


def adjoint_psd_full_f64(factor, rinv, a_svec, y, x_svec, c_svec=None, ybar=None, xbar=None, obj=None,
                         want_a=False):
    """Full adjoint at the central-path point (sdpx_gpu_adjoint_psd_full_f64):
    for functionals L_k = ybar_k . y + <Xbar_k, X> + obj_k <C, X> returns
    (bgrad m x r, cgrad s x r, agrad s x m x r or None)."""
    f64 = lambda a: None if a is None else np.asfortranarray(a, dtype=np.float64)
    rinv, a_svec, y, x_svec, c_svec = f64(rinv), f64(a_svec), f64(y), f64(x_svec), f64(c_svec)
    n = rinv.shape[0]
    s, m = a_svec.shape
    cols = [a.shape[1] for a in (ybar, xbar) if a is not None and np.ndim(a) == 2]
    r = cols[0] if cols else len(obj)
    ybar, xbar = f64(ybar), f64(xbar)
    obj = None if obj is None else np.ascontiguousarray(obj, dtype=np.float64)
    bgrad = np.zeros((m, r), order="F")
    cgrad = np.zeros((s, r), order="F")
    agrad = np.zeros((s, m, r), order="F") if want_a else None
    vp = ctypes.c_void_p
    ptr = lambda a: None if a is None else a.ctypes.data_as(vp)
    check(lib().sdpx_gpu_adjoint_psd_full_f64(factor._h, n, ptr(rinv), ptr(a_svec), ptr(y), ptr(x_svec),
                                              ptr(c_svec), r, ptr(ybar), ptr(xbar), ptr(obj), ptr(bgrad),
                                              ptr(cgrad), ptr(agrad)))
    return bgrad, cgrad, agrad


class PsdBlockF64(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("n", ctypes.c_uint64), ("ncols", ctypes.c_uint64),
                ("rinv", ctypes.c_void_p), ("a_svec", ctypes.c_void_p), ("columns", ctypes.c_void_p)]


def _condensed_dev(blocks, b_rows, w):
    """Device copies of a condensed KKT's PSD blocks and orthant rows: the
    sdpx_gpu_psd_block_f64 array, (b, w, rows) and the tensors to keep."""
    import torch

    dev = lambda a, dt=torch.float64: torch.tensor(np.asarray(a), dtype=dt, device="cuda")
    keep = []
    bl = (PsdBlockF64 * max(1, len(blocks)))()
    for i, (rinv, a_svec, cols) in enumerate(blocks):
        n = rinv.shape[0]
        r = dev(np.asfortranarray(rinv).T.reshape(-1))
        a = dev(np.asfortranarray(a_svec).T.reshape(-1))
        c = dev(np.asarray(cols, dtype=np.int64), torch.int64)
        keep += [r, a, c]
        bl[i] = PsdBlockF64(n, len(cols), r.data_ptr(), a.data_ptr(), c.data_ptr())
    bb = ww = None
    rows = 0
    if b_rows is not None:
        rows = b_rows.shape[0]
        bb, ww = dev(np.asfortranarray(b_rows).T.reshape(-1)), dev(w)
        keep += [bb, ww]
    return bl, bb, ww, rows, keep


def condensed_assemble_f64(m, p_csc, blocks, b_rows=None, w=None):
    """CondensedKKT::assemble on the GPU (sdpx_gpu_condensed_assemble_f64_dev):
    p_csc = (colptr, rowval, nzval) or None; blocks = [(rinv n x n,
    a_svec s x k, columns (k,))]; b_rows (r x m) and w (r,) the orthant
    rows.  Returns the m x m upper triangle (lower part zero)."""
    S = _condensed_assemble_dev(m, p_csc, blocks, b_rows, w)
    return S.cpu().numpy().reshape(m, m).T.copy()


def _condensed_assemble_dev(m, p_csc, blocks, b_rows, w):
    import torch

    dev = lambda a, dt=torch.float64: torch.tensor(np.asarray(a), dtype=dt, device="cuda")
    bl, bb, ww, rows, keep = _condensed_dev(blocks, b_rows, w)
    S = torch.zeros(m * m, dtype=torch.float64, device="cuda")
    pc = pr = pv = None
    if p_csc is not None:
        pc, pr, pv = dev(p_csc[0], torch.int64), dev(p_csc[1], torch.int64), dev(p_csc[2])
    ptr = lambda t: None if t is None else t.data_ptr()
    torch.cuda.synchronize()
    check(lib().sdpx_gpu_condensed_assemble_f64_dev(m, ptr(pc), ptr(pr), ptr(pv), len(blocks), bl, rows, ptr(bb),
                                                    max(rows, 1), ptr(ww), S.data_ptr(), m))
    return S


def condensed_solve_f64(m, p_csc, blocks, b_rows, w, bx, bz):
    """One condensed KKT solve on the GPU (CondensedKKT::solve_raw =
    prepare_rhs, the reduced solve, recover_rhs; kkt/condensed/mod.rs:
    2416-2610): S assembled and factored, xr = bx + A^T H^-1 bz, x = S^-1 xr,
    z = H^-1 (A x - bz).  bz and z: the blocks' svec rows, then the orthant
    rows.  Returns (x, z)."""
    import torch

    S = _condensed_assemble_dev(m, p_csc, blocks, b_rows, w)
    bl, bb, ww, rows, keep = _condensed_dev(blocks, b_rows, w)
    dbx = torch.tensor(np.asarray(bx, dtype=np.float64), device="cuda")
    dbz = torch.tensor(np.asarray(bz, dtype=np.float64), device="cuda")
    xr = torch.empty(m, dtype=torch.float64, device="cuda")
    z = torch.empty(dbz.numel(), dtype=torch.float64, device="cuda")
    ptr = lambda t: None if t is None else t.data_ptr()
    torch.cuda.synchronize()
    check(lib().sdpx_gpu_condensed_rhs_f64_dev(m, len(blocks), bl, rows, ptr(bb), max(rows, 1), ptr(ww),
                                               dbx.data_ptr(), dbz.data_ptr(), xr.data_ptr()))
    f = DenseFactor(m)
    assert f.refactor_dev(S.data_ptr(), m), "condensed Schur matrix not positive definite"
    f.solve_many_dev(xr.data_ptr(), 1, m)
    check(lib().sdpx_gpu_condensed_recover_f64_dev(m, len(blocks), bl, rows, ptr(bb), max(rows, 1), ptr(ww),
                                                   xr.data_ptr(), dbz.data_ptr(), z.data_ptr()))
    return xr.cpu().numpy(), z.cpu().numpy()


class ExactBlock(_Sized):
    _fields_ = [("struct_size", ctypes.c_uint64), ("rows", ctypes.c_uint64), ("ncols", ctypes.c_uint64), ("y", ctypes.c_void_p),
                ("d", ctypes.c_void_p), ("columns", ctypes.c_void_p)]


def xgemm_blocks_upper_exact(m, blocks, out):
    """XgemmScalar::xgemm_blocks_upper_exact (traits.rs:250-258): blocks are
    (y records rows x ncols column-major, d records (rows), columns (ascending
    uint64)).  Writes c[i + j m], i <= j; returns False on a decline."""
    arr = (ExactBlock * len(blocks))()
    keep = []
    for i, (y, d, cols) in enumerate(blocks):
        cols = np.ascontiguousarray(cols, dtype=np.uint64)
        keep.append(cols)
        arr[i] = ExactBlock(d.size, cols.size, _ptr(y), _ptr(d), _ptr(cols))
    nl = mp.limbs_of(out)
    code = lib().sdpx_gpu_rns_gemm_blocks_upper(nl, m, len(blocks), arr, _ptr(out))
    check(code, allow_decline=True)
    return code == OK


def xgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """XgemmScalar::xgemm for MPFR (blas/mpfr.rs:245-348): c = axpby(alpha,
    RN(op(a) op(b)), beta, c) with c of leading dimension ldc; alpha and beta
    are one-value record arrays.  False on a decline (c untouched)."""
    code = lib().sdpx_gpu_xgemm(mp.limbs_of(c), ord(ta), ord(tb), m, n, k, _ptr(alpha), _ptr(a), lda, _ptr(b), ldb,
                                _ptr(beta), _ptr(c), ldc)
    check(code, allow_decline=True)
    return code == OK


def xsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    """XsyrkScalar::xsyrk for MPFR (blas/mpfr.rs:540-626): the uplo triangle
    of c = axpby(alpha, RN(op(a) op(a)^T), beta, c); c column-packed upper
    when uplo is 'U', ldc == n and c has n(n+1)/2 values.  False on a
    decline (c untouched)."""
    code = lib().sdpx_gpu_xsyrk(mp.limbs_of(c), ord(uplo), ord(trans), n, k, _ptr(alpha), _ptr(a), lda,
                                _ptr(beta), _ptr(c), ldc, c.size)
    check(code, allow_decline=True)
    return code == OK


def diag_congruence_upper_exact(m, k, a, d, c):
    """XgemmScalar::diag_congruence_upper_exact (traits.rs:259-267,
    rns_blas.rs:2997-3195): the upper triangle of a^T diag(d) a (a k x m,
    ld k) into c[i + j m], each entry rounded once.  False on a decline."""
    code = lib().sdpx_gpu_diag_congruence_upper_exact(mp.limbs_of(c), m, k, _ptr(a), _ptr(d), _ptr(c))
    check(code, allow_decline=True)
    return code == OK


def xcongruence_exact(transa, m, k, a, lda, x, ldx, c, upper_only):
    """XgemmScalar::xcongruence_exact (traits.rs:290-296, rns_blas.rs:
    3563-3790): op(a) x op(a)^T into c (ldc = m), upper triangle only when
    upper_only.  False on a decline."""
    code = lib().sdpx_gpu_congruence_exact(mp.limbs_of(c), ord(transa), m, k, _ptr(a), lda, _ptr(x), ldx,
                                           _ptr(c), 1 if upper_only else 0)
    check(code, allow_decline=True)
    return code == OK


def xsvec_quadratic_exact(h, kmax, q, abs_q, x, sqrt2, out):
    """XgemmScalar::xsvec_quadratic_exact (traits.rs:272-282, rns_blas.rs:
    4032-4173): out[k] = sum_{i<=j} c_ij x_t q_ik q_jk (svec x, c_ij = sqrt2
    off the diagonal; |q| when abs_q).  sqrt2 is a one-value record array.
    False on a decline."""
    code = lib().sdpx_gpu_svec_quadratic_exact(mp.limbs_of(out), h, kmax, _ptr(q), 1 if abs_q else 0, _ptr(x),
                                               _ptr(sqrt2), _ptr(out))
    check(code, allow_decline=True)
    return code == OK


def xsymmetric_bilinear_exact(h, columns, q, x, pairs, out):
    """XgemmScalar::xsymmetric_bilinear_exact (traits.rs:283-289, rns_blas.rs:
    4175-4276): out[p] = q_a^T X q_b for pairs[p] = (a, b), X symmetric as
    its packed upper triangle.  False on a decline."""
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1))
    code = lib().sdpx_gpu_symmetric_bilinear_exact(mp.limbs_of(out), h, columns, _ptr(q), _ptr(x),
                                                   len(pairs), _ptr(pr), _ptr(out))
    check(code, allow_decline=True)
    return code == OK


def xpotrf(uplo, n, a, lda):
    """XpotrfScalar::xpotrf for MPFR (blas/mpfr.rs:727-789) on the GPU: a is
    a record array (column-major, leading dimension lda), factored in place.
    Returns info (0, the failing pivot column + 1, or -1/-4 for a bad uplo
    or lda); raises on a decline."""
    nl = mp.limbs_of(a)
    info = ctypes.c_int32(0)
    check(lib().sdpx_gpu_mp_potrf(nl, ord(uplo), n, _ptr(a), lda, ctypes.byref(info)))
    return info.value


def xpotrs(uplo, n, nrhs, a, lda, b, ldb):
    """XpotrsScalar::xpotrs for MPFR (blas/mpfr.rs:790-872): b (n x nrhs,
    leading dimension ldb) is overwritten by the solution.  Returns info."""
    nl = mp.limbs_of(a)
    info = ctypes.c_int32(0)
    check(lib().sdpx_gpu_mp_potrs(nl, ord(uplo), n, nrhs, _ptr(a), lda, _ptr(b), ldb,
                                  ctypes.byref(info)))
    return info.value


def xtrsm_lower(n, a, b, transpose):
    """xtrsm_lower for MPFR (blas/traits.rs:85-113): L X = B (or L^T X = B)
    with a n x n and b n x nrhs, both with leading dimension n; b in place."""
    nl = mp.limbs_of(a)
    nrhs = b.size // max(n, 1) if n else 0
    check(lib().sdpx_gpu_mp_trsm_lower(nl, n, _ptr(a), _ptr(b), nrhs, 1 if transpose else 0))


def mp_ldl_solve(l, dinv, n, b):
    """L D L^T x = b on the GPU (b: records, n x nrhs column-major).
    Returns x (new array); raises on a decline."""
    nl = mp.limbs_of(l)
    x = b.copy()
    nrhs = x.size // n
    check(lib().sdpx_gpu_mp_ldl_solve(nl, n, _ptr(l), n, _ptr(dinv), nrhs, _ptr(x), n))
    return x
