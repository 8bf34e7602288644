import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch ships its own HIP runtime under the same soname as ROCm's.  Loading
# torch before the search library makes both use torch's copy, so GPU tests
# can pass torch tensors and streams to the library; the other order leaves
# torch without a device.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(autouse=True)
def _oracle_faithful():
    """Every test starts with the oracle's planted-mistake switch off."""
    import oracle
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)
