import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "carmichael-frontier_amd")
for p in (PKG, os.path.join(PKG, "gpu"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcfjoin on the device)")
    config.addinivalue_line("markers", "slow: CPU tests over the full-size anchors")


@pytest.fixture(scope="session")
def workloads():
    from cfjoin import instance
    return instance.load_workloads()
