"""The allocation-free bandwidth sweep (kevd_bandwidth_sweep_noalloc): its
intermediates live in the library's fixed per-device scratch array instead of
context buffers.  It runs the same kernels with the same launch shapes as
kevd_bandwidth_sweep, so its sums must equal that entry point's bit for bit,
on consecutive calls of decreasing size (the array keeps the previous call's
contents) and from two contexts sharing the device's array."""
import numpy as np
import pytest

from conftest import clustered_points

import oracle as orc
from kernel_evd_amd import KevdError
from kernel_evd_amd._lib import KEVD_EINVAL

pytestmark = pytest.mark.gpu


def test_noalloc_sweep_equals_the_allocating_sweep():
    from kernel_evd_amd import Context
    n = 20000
    X = clustered_points(n, 64, seed=31)
    eps = orc.median_sqdist(X, seed=0, nsample=100001) * np.logspace(-3, 3, 64)
    rng = np.random.default_rng(32)
    with Context(0) as c, Context(0) as c2:
        c.set_points(X)
        c2.set_points(X[:5000])
        for prec in ("mixed", "fp64"):
            for nrows in (3000, 517):  # the second call smaller than the first
                rows = rng.choice(n, nrows, replace=False)
                print("noalloc sweep", prec, nrows, flush=True)
                a = c.bandwidth_sweep(rows, eps, prec, noalloc=True)
                b = c.bandwidth_sweep(rows, eps, prec)
                assert np.array_equal(a, b), (prec, nrows, np.max(np.abs(a - b) / b))
            # a second context on the same device uses the same array
            rows2 = rng.choice(5000, 211, replace=False)
            a2 = c2.bandwidth_sweep(rows2, eps[:40], prec, noalloc=True)
            assert np.array_equal(a2, c2.bandwidth_sweep(rows2, eps[:40], prec))
        # FP64 sums against the CPU oracle on a few rows
        rows = np.arange(0, n, 997)
        np.testing.assert_allclose(c.bandwidth_sweep(rows, eps, "fp64", noalloc=True),
                                   orc.bandwidth(X, rows, eps), rtol=1e-12)


def test_noalloc_sweep_rejects_work_beyond_the_scratch_array():
    from kernel_evd_amd import Context
    X = clustered_points(20000, 64, seed=33)
    with Context(0) as c:
        c.set_points(X)
        rows = np.tile(np.arange(20000), 4)  # 80,000 rows: about 42 MB of FP64 intermediates
        with pytest.raises(KevdError) as ei:
            c.bandwidth_sweep(rows, np.array([1.0, 10.0, 100.0]), "fp64", noalloc=True)
        assert ei.value.status == KEVD_EINVAL and "scratch" in str(ei.value)
        # the entry point still works after the refusal
        S = c.bandwidth_sweep(np.arange(100), np.array([1.0, 10.0, 100.0]), "fp64", noalloc=True)
        assert np.array_equal(S, c.bandwidth_sweep(np.arange(100), np.array([1.0, 10.0, 100.0]), "fp64"))
