import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def amd():
    """The MI355X library; a GPU test fails (never falls back) if it is missing."""
    from jvsip_amd.vsip import load
    return load()


@pytest.fixture(scope="session")
def ref():
    from oracle import pyoracle
    if not pyoracle.ref_available():
        pytest.skip("oracle/_ref/libjvsip_ref.so not built (needs the reference sources)")
    return pyoracle.ref()
