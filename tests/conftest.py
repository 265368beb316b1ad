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
    """One xc_ctx on device 0 for the GPU tests (fails loudly if the library or device is missing).

    torch (whose ROCm build carries its own HIP runtime) is initialised on the device first, as the
    benchmark does: the workload generators make their inputs with it, and it does not find the device
    once the library's runtime has opened it."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception as e:  # the library's own tests do not need torch
        print(f"torch device init skipped: {e}")
    from alignment_projects_amd import xcorr
    c = xcorr.Context(0)
    yield c
    c.close()
