import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def gpu():
    """Asserts a GPU is present (GPU tests must not pass on a fallback)."""
    from craton_bolt_amd.device import device_count

    n = device_count()
    assert n >= 1, "GPU test collected on a host without a usable GPU"
    print(f"[gpu] devices={n}", flush=True)
    return n
