"""Shared test set-up: import paths, the `gpu` marker, library handles."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "sph-multiphase_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsphx.so")


@pytest.fixture(scope="session")
def sphx_mod():
    import sphx
    sphx.lib()
    return sphx


@pytest.fixture(scope="session")
def gpu(sphx_mod):
    if sphx_mod.device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X (no CPU fallback)")
    return sphx_mod


@pytest.fixture(scope="session")
def orc():
    import orc as _orc
    _orc.lib()
    return _orc
