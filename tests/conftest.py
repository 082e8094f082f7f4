import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_sessionstart(session):
    """The warm-start tests put torch device memory beside the engine
    library's own HIP runtime.  torch's runtime has to start first: an engine
    created before it leaves torch without a device.  So a session that may
    run GPU tests starts it before any test."""
    mark = session.config.getoption("-m") or ""
    if "not gpu" in mark.replace("  ", " "):
        return
    try:
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except Exception:
        pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long statistical test")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a failed run still says which test ran
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
