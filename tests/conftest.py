import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU run (published-log reproductions)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (binaries built from the reference sources)")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("mccs")
