import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import mobileraytracer_amd as mrt

    return mrt.Context(0)
