import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "ref: needs the compiled reference under oracle/_ref")


@pytest.fixture(scope="session")
def ctx():
    import vite_amd
    c = vite_amd.Context(0)
    yield c
    c.close()
