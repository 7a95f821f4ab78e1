import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("apg-sgm_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import apg_amd
    c = apg_amd.Context(0)
    yield c
    c.close()
