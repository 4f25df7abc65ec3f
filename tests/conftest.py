import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "barnes-hut_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def progress(capsys):
    """print() past pytest's output capture: long GPU tests report each phase
    as it ends, so a run that is merely long never looks silent."""
    def _progress(msg):
        with capsys.disabled():
            print(msg, flush=True)
    return _progress
