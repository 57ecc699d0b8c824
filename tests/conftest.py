import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a crash log identifies it
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[case] %s" % request.node.nodeid, flush=True)
    yield
