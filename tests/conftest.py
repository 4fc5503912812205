"""Shared pytest configuration: markers and library fixtures."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_bind

    return oracle_bind.load()


@pytest.fixture(scope="session")
def amd():
    import celeritas_amd

    if not celeritas_amd.device_available():
        pytest.fail("gpu test selected but no HIP device is available")
    return celeritas_amd
