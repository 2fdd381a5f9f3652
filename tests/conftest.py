import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "pfsf-core_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP path")


@pytest.fixture(scope="session")
def cuda():
    """The HIP device.  GPU tests never fall back: no device is a failure."""
    import torch

    assert torch.cuda.is_available(), "gpu test run without a HIP device"
    from pfsf_amd import capi

    capi.load()  # raises if libpfsf.so is missing
    return torch.device("cuda:0")
