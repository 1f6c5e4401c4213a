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
    L.sdpx_gpu_last_stage_times.argtypes = [dp, u32]
    L.sdpx_gpu_last_stage_times.restype = i32
    gemm_args = [u32, u8, u8, u64, u64, u64, vp, u64, vp, u64, u32, vp, u64]
    for name in ("sdpx_gpu_rns_gemm", "sdpx_gpu_rns_gemm_dev"):
        f = getattr(L, name)
        f.argtypes = gemm_args
        f.restype = i32


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


def last_stage_times():
    """Stage times (ms) of this thread's last exact product: scan, encode,
    product, crt (needs set_profiling(True))."""
    buf = (ctypes.c_double * 4)()
    lib().sdpx_gpu_last_stage_times(buf, 4)
    return list(buf)


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
