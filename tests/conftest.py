import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    import harm_harm_amd as H
    return H.device_count() > 0


@pytest.fixture(scope="session")
def hlib():
    import harm_harm_amd as H
    return H.lib()


@pytest.fixture(scope="session")
def olib():
    import _oracle
    return _oracle.lib()
