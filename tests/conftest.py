import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ipdps15-pattern_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through libpattern's C ABI)")


@pytest.fixture(scope="session")
def pa():
    import pattern_amd
    pattern_amd.lib()
    return pattern_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(pa):
    """Initialises device 0 through the C ABI.  Skips only on a machine with
    no GPU driver at all; on a GPU machine a failing init is an error."""
    if not os.path.exists("/dev/kfd"):
        pytest.skip("no GPU driver on this machine")
    pa.init(0)
    return pa
