import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work starts, so a fault can be attributed
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
