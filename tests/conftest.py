import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def runtime():
    import unicorn_amd as U
    U.init(int(os.environ.get("UA_TEST_CPU_STUBS", "4")))
    return U


@pytest.fixture(scope="session")
def gpu_runtime(runtime):
    if not runtime.has_gpu():
        pytest.fail("GPU test selected but no HIP device is visible to the runtime")
    return runtime
