import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def fr3():
    from isaac_sim_ros2_humble_amd.model import load_robot
    return load_robot("fr3")


@pytest.fixture(scope="session")
def fetch():
    from isaac_sim_ros2_humble_amd.model import load_robot
    return load_robot("fetch")


@pytest.fixture(scope="session")
def orc_fr3(fr3):
    import oracle
    return oracle.OracleRobot(fr3)


@pytest.fixture(scope="session")
def orc_fetch(fetch):
    import oracle
    return oracle.OracleRobot(fetch)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
