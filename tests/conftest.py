import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def _have_gpu():
    try:
        from letkf_amd import _lib
        return _lib.lib().lekf_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail loudly when the device or the library is missing."""
    from letkf_amd import _lib
    n = _lib.lib().lekf_device_count()
    assert n > 0, "no GPU visible to liblekf_hip.so"
    return n


@pytest.fixture(scope="session")
def ref_lib():
    import oracle
    if not oracle.ref_available():
        pytest.skip("reference Fortran library oracle/_ref not built")
    return oracle
