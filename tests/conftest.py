import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session", autouse=True)
def _native_libraries():
    """Build whatever is missing (the build itself runs on the CPU)."""
    import hemepure_amd
    import oracle

    if not hemepure_amd.LIB_PATH.exists():
        from hemepure_amd import build

        build.build(verbose=False)
    if not oracle.LIB_PATH.exists():
        oracle.build()
    yield
