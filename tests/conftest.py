import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "simple-cpsd_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not gpu_present():
        pytest.skip("no GPU in this container")
    import cpsd_amd
    return cpsd_amd.hip_lib()
