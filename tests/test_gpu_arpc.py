"""GPU parity of the kernel block and ARPC against the CPU oracle.

Bars: FP64 kernel block 1e-13 relative; FP32 (mixed) 2e-6 absolute; pivots
bit-exact for FP64 with the same seed (and for mixed at these sizes); the
FP64 factor within 1e-9 of the oracle's; the final mixed trace error within
5% of FP64's.  The mixed factor on the oracle's injected pivots is compared
with the oracle's FP64 factor and with the rounding model
(oracle/fp32_model.py) of the same configuration: it must be within a
factor 2 of the model's prediction.  The model also shows (tests/
test_fp32_model.py) that no factor whose kernel values carry FP32 rounding
gets within 1e-6 of the FP64 factor.
"""
import numpy as np
import pytest

from conftest import clustered_points, gauss_points

import oracle as orc

pytestmark = pytest.mark.gpu


def _setup(ctx, X):
    ctx.set_points(X)
    print("points", X.shape, flush=True)


@pytest.mark.parametrize("dim", [3, 16, 64, 100])
def test_kernel_block_fp64(gpu_ctx, dim):
    X = gauss_points(300, dim, seed=dim)
    _setup(gpu_ctx, X)
    eps = orc.median_sqdist(X)
    rng = np.random.default_rng(1)
    rows = rng.integers(0, 300, 77)
    cols = np.concatenate([rng.integers(0, 300, 150), rows[:5]])
    K = gpu_ctx.kernel_block(rows, cols, eps, "fp64")
    R = orc.kernel_block(X, rows, cols, eps)
    np.testing.assert_allclose(K, R, rtol=1e-12, atol=1e-14)
    # diagonal pairs are exactly 1
    for i, r in enumerate(rows):
        for j, c in enumerate(cols):
            if r == c:
                assert K[i, j] == 1.0


@pytest.mark.parametrize("dim", [16, 64])
def test_kernel_block_mixed(gpu_ctx, dim):
    X = gauss_points(513, dim, seed=7)
    _setup(gpu_ctx, X)
    eps = orc.median_sqdist(X)
    rows = np.arange(513)
    cols = np.arange(0, 513, 3)
    K = gpu_ctx.kernel_block(rows, cols, eps, "mixed")
    R = orc.kernel_block(X, rows, cols, eps)
    assert np.max(np.abs(K - R)) < 2e-6
    assert np.all(K > 0) and np.all(K <= 1.0)


def test_median(gpu_ctx):
    X = gauss_points(1000, 8, seed=3)
    _setup(gpu_ctx, X)
    assert abs(gpu_ctx.median_sqdist() - orc.median_sqdist(X)) <= 1e-12 * orc.median_sqdist(X)
    a = gpu_ctx.median_sqdist(seed=5, nsample=20001)
    b = orc.median_sqdist(X, seed=5, nsample=20001)
    assert abs(a - b) <= 1e-12 * b


# block sizes 40 and 100 leave a partial 16-proposal group in the rejection
# block's register layout (16 + 16 + 8 and 6 x 16 + 4 proposals)
@pytest.mark.parametrize("block,rank", [(1, 40), (16, 100), (40, 150), (100, 250), (128, 300)])
def test_arpc_fp64_matches_oracle(gpu_ctx, block, rank):
    X = clustered_points(2000, 12, seed=block)
    _setup(gpu_ctx, X)
    eps = orc.median_sqdist(X)
    ref = orc.arpc(X, eps, rank, block=block, seed=11)
    out = gpu_ctx.arpc(eps, rank, block=block, seed=11, prec="fp64")
    print("rank", out.rank, "rounds", out.rounds, flush=True)
    assert out.rank == ref.rank
    np.testing.assert_array_equal(out.pivots, ref.pivots)
    np.testing.assert_array_equal(out.accepted, ref.accepted)
    np.testing.assert_allclose(out.trace_err, ref.trace_err, rtol=1e-9, atol=1e-14)
    F = gpu_ctx.get_factor(out.rank, "fp64")
    assert np.max(np.abs(F - ref.F)) < 1e-9
    # interpolation: F F(S,:)^T = K(:,S)
    KS = orc.kernel_block(X, np.arange(X.shape[0]), out.pivots, eps)
    assert np.max(np.abs(F @ F[out.pivots].T - KS)) < 1e-8


def test_arpc_mixed_factor_vs_oracle_and_rounding_model(gpu_ctx):
    """16,384 KS snapshots, rank 1000, block 128 (final trace error
    1.8e-4): the model predicts 6e-4 for the mixed factor's relative
    Frobenius distance from FP64 on the same pivots."""
    import fp32_model as fm
    from kernel_evd_amd import ks_generate
    n, rank = 16384, 1000
    X = ks_generate(n, L_dom=22.0, M=64, dt=0.25, n_discard=10000, stride=4, seed=0)
    eps = orc.median_sqdist(X, seed=0, nsample=1 << 20)
    ref = orc.arpc(X, eps, rank, block=128, seed=0)
    Fm, piv_m, _ = fm.factor(X, eps, ref.pivots, rank, block=128, variant="mixed")
    assert np.array_equal(piv_m, ref.pivots)
    model = fm.rel_frobenius(Fm, ref.F)
    _setup(gpu_ctx, X)
    out = gpu_ctx.arpc(eps, rank, block=128, seed=0, prec="mixed", inject_pivots=ref.pivots)
    assert out.rank == rank
    np.testing.assert_array_equal(out.pivots, ref.pivots)
    F = gpu_ctx.get_factor(rank, "fp64")
    rel = fm.rel_frobenius(F, ref.F)
    print("mixed factor vs oracle FP64: rel Frobenius %.3e (rounding model %.3e)" % (rel, model), flush=True)
    assert 0.25 * model < rel < 2.0 * model
    out64 = gpu_ctx.arpc(eps, rank, block=128, seed=0, prec="fp64", inject_pivots=ref.pivots)
    np.testing.assert_array_equal(out64.pivots, ref.pivots)
    assert fm.rel_frobenius(gpu_ctx.get_factor(rank, "fp64"), ref.F) < 1e-10


def test_arpc_mixed_pivots_and_factor(gpu_ctx):
    X = clustered_points(3000, 16, seed=4)
    _setup(gpu_ctx, X)
    eps = orc.median_sqdist(X)
    r64 = gpu_ctx.arpc(eps, 256, block=64, seed=3, prec="fp64")
    F64 = gpu_ctx.get_factor(r64.rank, "fp64")
    rmx = gpu_ctx.arpc(eps, 256, block=64, seed=3, prec="mixed", inject_pivots=r64.pivots)
    Fmx = gpu_ctx.get_factor(rmx.rank, "fp64")
    np.testing.assert_array_equal(rmx.pivots, r64.pivots)
    rel = np.linalg.norm(Fmx - F64) / np.linalg.norm(F64)
    print("mixed vs fp64 factor rel Frobenius", rel, flush=True)
    assert rel < 1e-4  # 3000 clustered points: trace error ~1e-3, so little amplification
    # trace error agrees with FP64 to within 5%
    assert abs(rmx.trace_err[-1] - r64.trace_err[-1]) <= 0.05 * r64.trace_err[-1]


def test_arpc_full_rank_reproduces_K(gpu_ctx):
    X = gauss_points(150, 3, seed=9)
    _setup(gpu_ctx, X)
    eps = 4.0 * orc.median_sqdist(X)
    out = gpu_ctx.arpc(eps, 150, block=32, seed=2, tol=0.0, prec="fp64")
    F = gpu_ctx.get_factor(out.rank, "fp64")
    K = orc.dense_kernel(X, eps)
    # residual diagonal bounds the error of a PSD approximation
    err = np.abs(K - F @ F.T).max()
    assert err <= 1.5 * out.trace_err[-1] * 150 + 1e-10


def test_linv_export(gpu_ctx):
    X = clustered_points(800, 6, seed=12)
    _setup(gpu_ctx, X)
    eps = orc.median_sqdist(X)
    out = gpu_ctx.arpc(eps, 60, block=16, seed=1, prec="fp64")
    Li = gpu_ctx.get_linv(out.rank)
    KSS = orc.kernel_block(X, out.pivots, out.pivots, eps)
    # K(S,S) = L L^T  =>  L^-1 K(S,S) L^-T = I
    M = Li @ KSS @ Li.T
    assert np.max(np.abs(M - np.eye(out.rank))) < 1e-6
    assert np.allclose(np.triu(Li, 1), 0.0)
