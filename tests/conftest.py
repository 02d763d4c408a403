import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhcs.so on the device)")


@pytest.fixture(scope="session")
def engine():
    from mujoco_contact_surfaces_amd import Engine
    e = Engine(0)
    yield e
    e.close()
