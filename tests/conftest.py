import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a host without a device fails loudly."""
    import fgkl_amd

    n = fgkl_amd.device_count()
    assert n > 0, "gpu test run without a visible HIP device"
    return n
