import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "askap-benchmarks_amd"), os.path.join(REPO, "oracle"), REPO,
          os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtcv.so on the device)")
    config.addinivalue_line("markers", "slow: full-size oracle comparisons (tens of seconds)")


@pytest.fixture(scope="session")
def ctx():
    import tcv
    c = tcv.Context(0)
    yield c
    c.close()


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any device work so a failing run can be traced
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu case] {request.node.nodeid}", flush=True)
    yield
