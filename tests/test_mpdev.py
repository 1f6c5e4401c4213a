"""The GPU panel arithmetic (sdpx.rs_amd/csrc/mpdev.h), compiled for the
host, agrees bit for bit with the host library's arithmetic (mpx.h) on
random values: products, sums, differences, reciprocals and exact dots
(with and without a subtrahend) at 64..512 bits; its division and square
root agree bit for bit with the C oracle's (both correctly rounded).  Both
are pinned further against the oracles through the LDL^T and Cholesky
tests."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mpdev_matches_host_arithmetic(tmp_path):
    exe = str(tmp_path / "mpdev_selftest")
    src = os.path.join(HERE, "native", "mpdev_selftest.cpp")
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not present")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", src, "-o", exe,
                    "-ldl"], check=True, capture_output=True)
    from helpers import oracle
    so = oracle()._name
    r = subprocess.run([exe, "6000", so], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "mismatches 0" in r.stdout
