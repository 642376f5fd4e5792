import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "distance-occlusion_amd")
for p in (os.path.join(ROOT, "tests"), PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle_bind import load_oracle

    return load_oracle()


@pytest.fixture(scope="session")
def reference():
    """The reference compiled from its sources, or None where not built."""
    from oracle_bind import load_reference

    return load_reference()


@pytest.fixture(scope="session")
def lib():
    import dolhip

    L = dolhip.load()
    return L


@pytest.fixture(scope="session")
def gpu(lib):
    """The library with a usable device; GPU tests fail (not skip) without one."""
    n = lib.device_count()
    assert n > 0, "no gfx950 device: " + lib.last_error()
    return lib


@pytest.fixture
def rng():
    return np.random.default_rng(20261203)
