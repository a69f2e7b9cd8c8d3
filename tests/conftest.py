"""pytest configuration: the `gpu` marker and shared fixtures.

Tests marked `gpu` need a gfx950 device and the in-tree libtrid.so; the CPU
suite (-m "not gpu") covers the oracle, the golden fixtures, the host logic
and the C ABI's argument validation without touching a device.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X) and libtrid.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def trid():
    import tridiagonalsolvers_amd as T
    T.lib()  # raises if libtrid.so is missing: no CPU fallback exists
    return T
