import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (runs the device library)")


@pytest.fixture(scope="session")
def oracle_built():
    import pyoracle
    if not os.path.exists(pyoracle.ORACLE_LIB):
        pyoracle.build(ref=False)
    return pyoracle
