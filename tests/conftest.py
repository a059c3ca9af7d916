"""Shared pytest setup: import paths, the `gpu` marker, seeded helpers."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PARENT = os.path.join(REPO, "hpx-fft_amd")
for p in (REPO, PKG_PARENT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: full-size cases (seconds each)")


@pytest.fixture(scope="session")
def gpu():
    """torch on a HIP device; fails loudly (never skips) when there is none."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    import hpxfft_amd
    hpxfft_amd.lib()   # raises if libhpxfft_hip.so is missing
    return torch.device("cuda:0")
