import os
import sys

import pytest

# torch-ROCm ships its own HIP runtime; it must be loaded before the library
# (which links the system one) or torch finds no devices in a process where
# the library initialised HIP first.  Tests that use torch as the caller's
# allocator (device buffers for the pointer entry points) rely on this.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    from ccsd_parallel_amd import capi
    return capi.load()
