import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gpu-inner-outer_amd", "python"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU run (large oracle cases)")


@pytest.fixture(scope="session")
def oracle():
    import gsc_oracle

    gsc_oracle.load()
    return gsc_oracle


@pytest.fixture(scope="session")
def gpu():
    """The product library on a real device; fails (not skips) if the library
    is missing, since a gpu-marked test must run the HIP path."""
    from gsc_amd import capi

    lib = capi.load()
    if lib.gsc_device_count() <= 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return lib
