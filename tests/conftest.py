"""Shared fixtures; the ``gpu`` marker selects tests that need an MI355X."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libbhah_hip.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on CPU")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nrpy_amd import runtime

    runtime.load_library()
    return runtime
