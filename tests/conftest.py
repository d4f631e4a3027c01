"""Shared pytest configuration.

Markers: ``gpu`` marks a test that needs an MI355X (run with ``-m gpu``); every
other test runs on the CPU.  GPU tests fail (not skip) when the HIP engine is
missing, so a silent fallback cannot pass them.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built engine")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def cref():
    import oracle_lib
    return oracle_lib.cref()


@pytest.fixture(scope="session")
def ref():
    """The real reference CpuBackend, or skip where it could not be built."""
    import oracle_lib
    r = oracle_lib.ref()
    if r is None:
        pytest.skip("reference CPU sources not available to build oracle/_ref")
    return r


@pytest.fixture(scope="session")
def engine():
    """The HIP engine's device-0 context (GPU tests only)."""
    from steppe_amd import _core
    return _core.Context.get(0)
