"""pytest configuration: the `gpu` marker and shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def oracle_portable():
    from oracle.oracle import Oracle
    return Oracle("portable")


@pytest.fixture(scope="session")
def oracle_libm():
    from oracle.oracle import Oracle
    return Oracle("libm")


@pytest.fixture(scope="session")
def backend():
    """The product's Python API over liboptphot_hip.so (GPU tests only)."""
    from gpu_optical_photons_amd import backend as b
    if not b.gpu_available():
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X")
    return b
