import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built library")
    config.addinivalue_line("markers", "slow: long-running (full-size configurations)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle as po
    po.lib()
    return po
