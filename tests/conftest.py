"""Shared test setup: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "gpu-domain-propagator_amd")
for p in (PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle_binding
    return oracle_binding


@pytest.fixture(scope="session")
def gdp():
    import gdp_amd
    gdp_amd.load_library()
    return gdp_amd


@pytest.fixture(scope="session")
def gpu(gdp):
    # The device is probed through the library's own HIP runtime: importing a
    # framework that bundles another ROCm runtime into this process would
    # break device discovery for one of the two.
    if gdp.interface.device_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    return gdp
