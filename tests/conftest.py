import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X, gfx950)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The engine library on a visible GPU; GPU tests fail (not skip) when it is missing."""
    from tecs_amd import _lib

    lib = _lib.load()
    if _lib.device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return lib
