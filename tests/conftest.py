import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of libhipsw")


@pytest.fixture(scope="session")
def engine():
    from swiftbwa_amd.metalsw import HipSWEngine
    eng = HipSWEngine(0)          # raises loudly when the library or device is missing
    yield eng
    eng.close()
