import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libxcorr.so")


@pytest.fixture(scope="session")
def ctx():
    """One xc_ctx on device 0 for the GPU tests (fails loudly if the library or device is missing)."""
    from alignment_projects_amd import xcorr
    c = xcorr.Context(0)
    yield c
    c.close()
