import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()  # builds liboracle.so from the committed Makefile if absent
    return O


@pytest.fixture(scope="session")
def dam_params():
    from akuaengine_amd.config import dam_break_params
    return dam_break_params(4)
