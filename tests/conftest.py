"""Test configuration: the `gpu` marker (tests that need the MI355X) and import paths for the
package (rum-testing-gpu_amd/solver) and the CPU oracle (oracle/)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rum-testing-gpu_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (librumgpu kernels run)")


def pytest_collection_modifyitems(config, items):
    # print the test id before each GPU test runs, so a run that dies names its case
    pass


@pytest.fixture(autouse=True)
def _name_the_case(request):
    if request.node.get_closest_marker("gpu"):
        print("\n[gpu case] %s" % request.node.nodeid, flush=True)
    yield
