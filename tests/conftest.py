import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libkevd.so")


def gauss_points(n, dim, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal((n, dim))


def clustered_points(n, dim, seed=0):
    """Points on a few anisotropic clusters: a kernel with decaying spectrum."""
    rng = np.random.default_rng(seed)
    centers = 3.0 * rng.standard_normal((6, dim))
    lab = rng.integers(0, 6, n)
    scales = np.linspace(0.3, 1.0, dim)
    return centers[lab] + rng.standard_normal((n, dim)) * scales


@pytest.fixture(scope="session")
def gpu_ctx():
    from kernel_evd_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
