import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ctx():
    """One femgpu context for the GPU tests of a session."""
    import cudafemproject_amd as F
    c = F.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def fem_main():
    """Path of the FEMproject-compatible driver, rebuilt (g++, seconds) when it
    is missing or older than its sources."""
    from cudafemproject_amd.build import ensure_fem_main
    return ensure_fem_main()
