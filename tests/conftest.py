import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test on a machine without a ROCm GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def dev():
    return gpu_device()
