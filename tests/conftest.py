import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU-side cases")


def _gpu_available():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One libqbt context for the GPU tests of the session."""
    if not _gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
    import qubatron_amd as q
    ctx = q.Context(0)
    yield ctx
    ctx.close()
