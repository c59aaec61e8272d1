import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running (minutes)")


@pytest.fixture(scope="session")
def dev():
    """A device context on GPU 0 (fails, does not skip, when the HIP library
    or the GPU is missing: -m gpu tests only run on a GPU host)."""
    from sep_code_amd.device import BornDevice
    d = BornDevice(0)
    yield d
    d.close()
