import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long CPU test")
    # Load the library (and the HIP runtime it was built against) before any
    # test module imports torch, which ships its own copy of that runtime
    # under the same soname: the first one loaded serves the process.
    try:
        import cpic_amd
        cpic_amd.lib()
    except (OSError, RuntimeError):
        pass   # not built: the tests that need it fail on their own


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")
