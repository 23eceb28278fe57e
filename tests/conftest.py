import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    from sparrow_amd import _native
    c = _native.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracles():
    from oracle import oracle as O
    return {m: O.Oracle(m) for m in ("MNDO", "AM1", "PM3", "RM1", "PM6")}
