import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "yang-mills_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libym_amd.so")
    config.addinivalue_line("markers", "slow: CPU test that runs for more than a few seconds")


def gpu_available():
    """True when a HIP device node is visible (no HIP runtime is initialised)."""
    return os.path.exists("/dev/kfd") and any(n.startswith("renderD") for n in os.listdir("/dev/dri")) \
        if os.path.isdir("/dev/dri") else False


@pytest.fixture(scope="session")
def ym():
    import ym_amd
    ym_amd.lib()
    return ym_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib
