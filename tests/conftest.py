import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def cuda():
    """The HIP device. GPU tests fail (not skip) without one: they are only selected where a
    GPU is expected."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible to a test marked gpu")
    return torch.device("cuda:0")
