import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before it runs, so a run that dies on the device says where
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
