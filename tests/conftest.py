import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail (not skip) when it is missing."""
    from molmodel_amd import engine
    L = engine.lib()
    assert L.mm_device_count() > 0, "no HIP device visible"
    return L
