import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import utopia_amd as ua
    if ua.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    c = ua.Context(0)
    yield c
    c.close()


@pytest.fixture
def golden():
    return GOLDEN
