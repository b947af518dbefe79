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


@pytest.fixture(scope="session")
def elasticity_17():
    """P1 elasticity on the 17^3 Kuhn cube (3x3 blocks) with a load on the
    free vectors."""
    import numpy as np
    from ug_amd import problems
    P = problems.elasticity(4)
    lv = P.levels[-1]
    rhs = np.random.default_rng(0x5EED0801).uniform(-1, 1, 3 * lv.n)
    rhs[np.repeat(lv.skip != 0, 3)] = 0.0
    P.rhs = rhs
    return P
