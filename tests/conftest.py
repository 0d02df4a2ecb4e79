import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built libflusi_hip.so")
    config.addinivalue_line("markers", "slow: oracle runs of whole reference fixtures (tens of seconds)")


@pytest.fixture
def golden():
    return GOLDEN
