import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fouriercontourcells.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfcc_hip.so)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_engine():
    import oracle
    return oracle.engine()


@pytest.fixture(scope="session")
def hip():
    import fcc_hip
    return fcc_hip.hip_engine()


def deformed_coef(seed=3, maxmode=6, amplitude=0.04):
    """deformed_coef (test/runtests.jl:17-23) with a numpy generator."""
    rng = np.random.default_rng(seed)
    c = np.zeros(2 * maxmode - 1)
    c[1:] = amplitude * (2 * rng.random(2 * maxmode - 2) - 1)
    c[0] = np.sqrt(1 - np.sum(c[1:] ** 2) / 2)
    return c
