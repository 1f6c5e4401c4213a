"""The C-ABI library builds, loads and exports every declared symbol
(no compute calls: this container has no GPU)."""
import os
import re

import sdpx_gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(sdpx_gpu_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = sdpx_gpu.lib()
    names = declared()
    assert len(names) >= 6
    for name in names:
        assert hasattr(L, name), name


def test_abi_version_and_error_channel():
    L = sdpx_gpu.lib()
    assert L.sdpx_gpu_abi_version() == 5
    # invalid input is rejected before any device work
    from sdpx_gpu import mp
    a = mp.zeros(4, 4)
    code = L.sdpx_gpu_rns_gemm(4, ord("X"), ord("N"), 2, 2, 2, sdpx_gpu._ptr(a), 2,
                               sdpx_gpu._ptr(a), 2, 0, sdpx_gpu._ptr(a), 4)
    assert code == sdpx_gpu.INVALID
    assert "transpose" in sdpx_gpu.last_error()


def test_cholesky_argument_codes_before_device_work():
    # xpotrf / xpotrs argument codes (blas/mpfr.rs:728-736, 801-813) come
    # back before any device call
    import ctypes
    from sdpx_gpu import mp
    L = sdpx_gpu.lib()
    a = mp.zeros(4, 2)
    info = ctypes.c_int32(99)
    assert L.sdpx_gpu_mp_potrf(2, ord("X"), 2, sdpx_gpu._ptr(a), 2, ctypes.byref(info)) == 0
    assert info.value == -1
    assert L.sdpx_gpu_mp_potrf(2, ord("L"), 2, sdpx_gpu._ptr(a), 1, ctypes.byref(info)) == 0
    assert info.value == -4
    assert L.sdpx_gpu_mp_potrs(2, ord("U"), 2, 1, sdpx_gpu._ptr(a), 2, sdpx_gpu._ptr(a), 1,
                               ctypes.byref(info)) == 0
    assert info.value == -7
    assert L.sdpx_gpu_mp_potrf(2, ord("L"), 0, None, 1, ctypes.byref(info)) == 0
    assert info.value == 0


def test_tuning_switch_names():
    L = sdpx_gpu.lib()
    assert L.sdpx_gpu_set_tuning(b"no_such_switch", 1) == sdpx_gpu.INVALID
    assert "tuning" in sdpx_gpu.last_error()
    assert L.sdpx_gpu_set_tuning(b"trsm_rl_room", 128) == sdpx_gpu.OK
