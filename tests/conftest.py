import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libitomo.so")


def _say(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work, so a lost GPU run still says what it was doing
    if request.node.get_closest_marker("gpu") is not None:
        _say("[gpu-test] %s" % request.node.nodeid)
    yield
