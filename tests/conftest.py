import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libugmg's C ABI)")


@pytest.fixture(scope="session")
def poisson2d():
    from ug_amd import problems
    return problems.poisson2d_129()


@pytest.fixture(scope="session")
def jump3d_33():
    from ug_amd import problems
    return problems.jump3d(5)


@pytest.fixture(scope="session")
def jump3d_17():
    from ug_amd import problems
    return problems.jump3d(4)
