import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcunfft_amd.so)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The product library on a real GPU; fails (never skips) when either is missing."""
    import torch

    import cunfft_amd

    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    cunfft_amd.lib()
    return cunfft_amd
