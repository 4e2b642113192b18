import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PACKAGE = ROOT / "custom-ray-tracing-shader_amd"
for p in (str(ROOT), str(PACKAGE), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: long CPU-oracle case")


@pytest.fixture(scope="session")
def lib():
    import native_api
    return native_api.load_library()
