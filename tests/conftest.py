"""Test configuration: the ``gpu`` marker and import paths.

``-m gpu`` tests need a HIP device and the built libmsc_slab.so; every other
test runs on the CPU (oracle pins, host logic, the C ABI's host-only entry
points and its exported symbols).
"""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libmsc_slab.so")


@pytest.fixture(scope="session")
def hip_device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a HIP device")
    return torch.device("cuda", 0)
