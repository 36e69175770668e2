import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def sw():
    import skyweaver_amd
    skyweaver_amd.lib()
    return skyweaver_amd


@pytest.fixture(scope="session")
def dev(sw):
    """The GPU, through the library's own HIP runtime (no torch)."""
    if sw.device_count() < 1:
        pytest.fail("GPU test selected but no GPU is visible")
    return 0
