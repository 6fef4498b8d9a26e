import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO / "ray-tracing_amd", REPO / "tests", REPO):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblinrt_hip.so on the device)")
    config.addinivalue_line("markers", "slow: minutes of CPU work")


@pytest.fixture(scope="session")
def has_gpu():
    import linrt
    return linrt.device_count() > 0
