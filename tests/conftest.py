"""Shared test setup: import paths, the `gpu` marker, oracle and engine fixtures."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "sph-particle-simulator_amd")
for p in (PKG, os.path.join(REPO, "oracle"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import sph_oracle

    sph_oracle.load()
    return sph_oracle


@pytest.fixture(scope="session")
def sphk():
    import sphk as mod

    mod.load()
    return mod


@pytest.fixture(scope="session")
def scenes():
    import sph_scenes

    return sph_scenes


R2 = np.float32(0.04) * np.float32(0.04)
