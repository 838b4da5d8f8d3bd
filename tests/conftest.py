import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP product path")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when the HIP device or libvhm.so is missing (no fallback)."""
    import voxelhashmt_amd as vhm
    ok, msg = vhm.device_available()
    if not ok:
        pytest.fail(f"GPU test needs a HIP device: {msg}")
    print(f"[gpu] {msg}", flush=True)
    return vhm
