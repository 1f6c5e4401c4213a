import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "sdpx.rs_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu():
    import sdpx_gpu
    import torch  # noqa: F401  (device presence check only)

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    sdpx_gpu.check(sdpx_gpu.lib().sdpx_gpu_init(0))
    return sdpx_gpu
