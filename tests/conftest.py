import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (select with -m gpu)")


@pytest.fixture(scope="session")
def ctx():
    """One galp_ctx on device 0 for every GPU test of the session."""
    from galp_amd import _abi
    c = _abi.Context(0)
    yield c
    c.close()
