import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def sr():
    return entry.load_package()


@pytest.fixture(scope="session")
def orc():
    return entry.load_oracle()


@pytest.fixture(scope="session")
def scenes(sr):
    import swraster_viewer_amd.scenes as sc
    return sc


@pytest.fixture(scope="session")
def env(sr):
    return sr.Environment.load()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)
