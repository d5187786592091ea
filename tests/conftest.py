import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "a-dda_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def gpu_available():
    try:
        import adda_amd
        return adda_amd.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hipdev():
    import adda_amd
    if adda_amd.device_count() == 0:
        pytest.fail("no HIP device visible for a test marked gpu")
    return adda_amd
