import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfilament_hip.so)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def pkg():
    import _helpers
    return _helpers.load_pkg()


@pytest.fixture(scope="session")
def orc():
    import _helpers
    return _helpers.load_oracle()
