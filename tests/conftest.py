import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built on demand; test infrastructure only)."""
    from oracle import oracle as O
    if not (os.path.exists(O.LIB_PATH) and os.path.exists(O.LIB_O3_PATH)):
        O.build(ref=False)
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (one process on the card)."""
    from nbody_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()
