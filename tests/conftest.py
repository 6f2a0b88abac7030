import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def pu():
    from helpers import load_package

    return load_package()


@pytest.fixture(scope="session")
def oracle():
    from helpers import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def gpu(pu):
    """The product library on a visible GPU; GPU tests fail loudly without one."""
    n = pu.device_count()
    assert n >= 1, "no HIP device visible: GPU tests need the MI355X"
    return pu
