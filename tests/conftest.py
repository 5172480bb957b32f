import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    """True when a HIP device is visible to the solver library."""
    try:
        import eikonalxx_amd as ekx
        return ekx.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ekx():
    import eikonalxx_amd as E
    if E.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return E
