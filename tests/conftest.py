import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP engine")


@pytest.fixture(scope="session")
def p32():
    import yatfhe_amd as yg
    return yg.PARAMS_T32_N1024


@pytest.fixture(scope="session")
def keys32(p32):
    import yatfhe_amd as yg
    return yg.keygen(p32, seed=0)


@pytest.fixture(scope="session")
def p56():
    import yatfhe_amd as yg
    return yg.PARAMS_T56_N2048


@pytest.fixture(scope="session")
def keys56(p56):
    import yatfhe_amd as yg
    return yg.keygen(p56, seed=0)
