import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libpgn.so")


@pytest.fixture(scope="session")
def ctx():
    import physim_amd as pg
    c = pg.Context(0)
    yield c
    c.close()
