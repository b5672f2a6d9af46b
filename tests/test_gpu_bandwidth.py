"""GPU parity of the fused bandwidth sweep against the FP64 CPU oracle
(oracle.bandwidth: brute-force exp over every (row, column) pair).
Bars: FP64 path 1e-12 relative; mixed path 2e-6 relative on every S(eps)
and 1e-4 absolute on the log-log slope."""
import numpy as np
import pytest

from conftest import clustered_points, gauss_points

import oracle as orc
from kernel_evd_amd import bandwidth_post

pytestmark = pytest.mark.gpu


def _grid(med, lo=1e-3, hi=1e3, k=64):
    return med * np.logspace(np.log10(lo), np.log10(hi), k)


@pytest.mark.parametrize("n,dim,nrows", [(1000, 8, 130), (3001, 64, 257), (517, 3, 517), (700, 100, 64), (1500, 40, 200)])
def test_sweep_matches_oracle(gpu_ctx, n, dim, nrows):
    X = clustered_points(n, dim, seed=n)
    gpu_ctx.set_points(X)
    med = orc.median_sqdist(X)
    rows = np.random.default_rng(1).choice(n, nrows, replace=False)
    eps = _grid(med)
    ref = orc.bandwidth(X, rows, eps)
    S64 = gpu_ctx.bandwidth_sweep(rows, eps, prec="fp64")
    np.testing.assert_allclose(S64, ref, rtol=1e-12)
    S32 = gpu_ctx.bandwidth_sweep(rows, eps, prec="mixed")
    rel = np.abs(S32 - ref) / ref
    print("mixed max rel", rel.max(), flush=True)
    assert rel.max() < 2e-6
    _, s_ref, _, _ = bandwidth_post(eps, ref)
    _, s32, eopt, mdim = bandwidth_post(eps, S32)
    assert np.max(np.abs(s32 - s_ref)) < 1e-4
    # limits: tiny eps keeps only the self pairs, huge eps counts every pair
    S = gpu_ctx.bandwidth_sweep(rows, np.array([1e-9 * med, 1e9 * med]), prec="mixed")
    assert abs(S[0] - nrows) < 1e-6 * nrows
    assert abs(S[1] - nrows * n) < 1e-5 * nrows * n


def test_sweep_gaussian_cloud_slope(gpu_ctx):
    """iid N(0, I_m): E K = (eps / (eps + 4))^(m/2) off the diagonal, so the
    log-log slope tends to m/2 * 4/(eps+4) ... -> 0 as eps grows and the
    slope at small eps approaches m/2 (the reference's mdim = 2 slope ~ m)."""
    m = 4
    X = gauss_points(4000, m, seed=2)
    gpu_ctx.set_points(X)
    rows = np.arange(0, 4000, 5)
    eps = np.logspace(-1, 3, 41)
    S = gpu_ctx.bandwidth_sweep(rows, eps, prec="mixed")
    Soff = S - rows.size
    E = rows.size * (4000 - 1) * (eps / (eps + 4.0)) ** (m / 2)
    assert np.max(np.abs(Soff / E - 1.0)[10:]) < 0.05
