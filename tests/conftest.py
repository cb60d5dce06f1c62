"""Shared fixtures: the library binding, the oracle, and the `gpu` marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bubble-simulation_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def bs():
    import bubblesim
    bubblesim.lib()
    return bubblesim


@pytest.fixture(scope="session")
def orc():
    from oracle import cpu as orc_mod
    orc_mod.lib()
    return orc_mod


@pytest.fixture(scope="session")
def gpu_available(bs):
    n = bs.device_count()
    if n == 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return n
