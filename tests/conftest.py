import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "recblock-sptrsv_amd")
for p in (REPO, PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (gfx950) and librecblock.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O
