import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librayon_hip.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    import rayon_amd as ra
    c = ra.Context(0)
    yield c
    c.close()
