import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP engine")
    config.addinivalue_line("markers", "slow: long-running restatement of a reference test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()
