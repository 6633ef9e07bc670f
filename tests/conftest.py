import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libskewengine.so)")


@pytest.fixture(autouse=True)
def _announce_gpu_test(request):
    # Name each GPU test before it runs so that the last printed line of a
    # run that dies on the device identifies the case.
    if request.node.get_closest_marker("gpu") is not None:
        sys.stdout.write("\n[gpu-test] %s\n" % request.node.nodeid)
        sys.stdout.flush()
    yield
