"""Test configuration: the `gpu` marker and shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (MI355X)")


@pytest.fixture(scope="session")
def rt():
    from gpugi_amd import rt as _rt
    _rt.lib()  # fails loudly when libgpugi_rt.so is not built
    return _rt


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as _o
    _o.lib()
    return _o


@pytest.fixture(scope="session")
def ctx(rt):
    c = rt.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def cornell(rt):
    return rt.HostScene.cornell()


@pytest.fixture(scope="session")
def hf64(rt):
    from gpugi_amd import workloads
    return rt.HostScene.heightfield(64, workloads.MASTER_SEED)
