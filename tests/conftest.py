"""Shared pytest setup.

``gpu`` marks tests that need an MI355X; the CPU suite runs with
``-m "not gpu"``.  GPU tests skip when no device is present unless
``PLAQUETTE_REQUIRE_GPU`` is set, the same skip-versus-require rule as the
reference (src/device.rs:793-811).
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def gpu():
    import plaquette_amd as P

    if not P.gpu_available():
        if os.environ.get("PLAQUETTE_REQUIRE_GPU"):
            pytest.fail("PLAQUETTE_REQUIRE_GPU is set but no HIP device is available")
        pytest.skip("no HIP device")
    return P


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
