import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(autouse=True, scope="session")
def oracle_threads():
    # The checker's results do not depend on its thread count; threads only make
    # the large frames finish sooner.
    import oracle
    oracle.set_threads(min(8, len(os.sched_getaffinity(0))))
    yield


@pytest.fixture(scope="session")
def ctx():
    import tileink_amd as tk
    return tk.NativeContext(0)
