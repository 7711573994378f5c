import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _announce(item):
    # Name the case before any GPU work so a fault leaves a trace.
    if item.get_closest_marker("gpu") is not None:
        sys.stderr.write(f"\n[gpu-test] {item.nodeid}\n")
        sys.stderr.flush()


def pytest_runtest_setup(item):
    _announce(item)


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")
