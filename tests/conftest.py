import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def gpu_visible():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("clay")


@pytest.fixture(scope="session")
def synth_root():
    """Shared directory for the (deterministic) synthetic instruments."""
    path = os.environ.get("CLAY_SYNTH_DIR", "/tmp/clay_synth")
    os.makedirs(path, exist_ok=True)
    return path
