import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / 'mi3-gpu_amd'))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def have_gpu():
    try:
        from mi3gpu import _lib
        return _lib.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope='session')
def gpu_ok():
    if not have_gpu():
        pytest.fail("no HIP device visible (or libmi3hip.so missing) for a gpu-marked test")
    return True


@pytest.fixture(scope='session')
def tmpout(tmp_path_factory):
    return tmp_path_factory.mktemp('mi3out')
