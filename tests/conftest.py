import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session."""
    from lambrex_amd import lbx
    c = lbx.Context(int(os.environ.get("LBX_DEVICE", "0")))
    yield c
    c.close()


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the running test on stdout before any GPU work, so a run that ends
    # the process still says which case it was in.
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
