import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "tests"))
sys.path.insert(0, str(REPO / "zikkurat-algebra_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950) and the built library")
    config.addinivalue_line("markers", "slow: large sizes")


def gpu_available():
    try:
        import zkgpu
        zkgpu.device_info()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def zk():
    """The GPU library; a GPU test fails loudly (not skips) when it cannot be loaded."""
    import zkgpu
    zkgpu.lib()
    zkgpu.device_info()
    return zkgpu
