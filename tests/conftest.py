import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.binding import Oracle, build

    build(ref=True)
    return Oracle()


@pytest.fixture(scope="session")
def quest():
    from dgquest_amd import quest as Q

    return Q.load(2)


@pytest.fixture(scope="session")
def quest32():
    from dgquest_amd import quest as Q

    return Q.load(1)


@pytest.fixture(scope="session")
def env(quest):
    e = quest.createQuESTEnv()
    yield e


@pytest.fixture(scope="session")
def env32(quest32):
    e = quest32.createQuESTEnv()
    yield e
