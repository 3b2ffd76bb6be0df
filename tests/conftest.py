import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "eulerian-liquid_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through the C ABI of libelq.so)")


@pytest.fixture(scope="session")
def ctx():
    import elq
    c = elq.Context(0)
    yield c
    c.close()
