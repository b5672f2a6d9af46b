"""GPU parity of the bistochastic kEVD.

Oracles: oracle.lowrank_evd (the F-form of ks_kevd.cu in FP64 NumPy),
oracle.dense_reference_evd (ks_reference.py:21-40) for the full-rank case,
oracle.lowrank_sinkhorn_evd for the Sinkhorn mode.  Bars: FP64 path 1e-9
relative on eigenvalues; mixed path top-32 within 1e-5 relative of FP64;
full rank against the dense oracle 1e-8; lambda_max = 1.
"""
import numpy as np
import pytest

from conftest import clustered_points, gauss_points

import oracle as orc

pytestmark = pytest.mark.gpu


def _factor(ctx, n=1500, dim=8, rank=200, seed=0):
    X = clustered_points(n, dim, seed=seed)
    ctx.set_points(X)
    eps = orc.median_sqdist(X)
    out = ctx.arpc(eps, rank, block=32, seed=seed, prec="fp64")
    return X, eps, out, ctx.get_factor(out.rank, "fp64")


def _align(U, V):
    s = np.sign(np.sum(U * V, axis=0))
    s[s == 0] = 1
    return U * s


def test_evd_fp64_matches_lowrank_oracle(gpu_ctx):
    X, eps, out, F = _factor(gpu_ctx)
    k = 24
    res = gpu_ctx.evd(out.rank, k_out=k, prec="fp64")
    lam, U, dev, shift = orc.lowrank_evd(F, k)
    print("lam top", res.lam[-4:], "ref", lam[-4:], flush=True)
    np.testing.assert_allclose(res.lam[-k:], lam[-k:], rtol=1e-9, atol=1e-13)
    assert abs(res.lam[-1] - 1.0) < 1e-10
    assert res.rowsum_dev < 1e-9
    assert abs(res.shift - shift) <= 1e-12 * shift
    # leading eigenvectors up to sign (skip near-degenerate ones)
    Ua = _align(res.U.astype(np.float64), U)
    gaps = np.abs(np.diff(lam[-k - 1:]))
    ok = (gaps[:-1] > 1e-6) & (gaps[1:] > 1e-6) if k > 1 else np.array([True])
    err = np.max(np.abs(Ua - U), axis=0)
    assert np.all(err[-8:][ok[-8:]] < 1e-5), err[-8:]


def test_evd_mixed_vs_fp64(gpu_ctx):
    X, eps, out, F = _factor(gpu_ctx, seed=1)
    k = 32
    r64 = gpu_ctx.evd(out.rank, k_out=k, prec="fp64")
    gpu_ctx.set_factor(F.astype(np.float32))
    rmx = gpu_ctx.evd(out.rank, k_out=k, prec="mixed")
    rel = np.abs(rmx.lam[-32:] - r64.lam[-32:]) / np.abs(r64.lam[-32:])
    print("mixed top-32 max rel err", rel.max(), flush=True)
    assert rel.max() < 1e-5
    # against the FP64 oracle on the same (FP32-rounded) factor
    lam, _, _, _ = orc.lowrank_evd(F.astype(np.float32).astype(np.float64), k)
    assert np.max(np.abs(rmx.lam[-32:] - lam[-32:]) / np.abs(lam[-32:])) < 1e-5


def test_full_rank_equals_dense_reference(gpu_ctx):
    n = 240
    X = gauss_points(n, 4, seed=5)
    gpu_ctx.set_points(X)
    eps = 2.0 * orc.median_sqdist(X)
    out = gpu_ctx.arpc(eps, n, block=16, seed=3, tol=0.0, prec="fp64")
    assert out.rank == n or out.trace_err[-1] < 1e-14
    res = gpu_ctx.evd(out.rank, k_out=32, prec="fp64", all_eigenvalues=True)
    K = orc.dense_kernel(X, eps)
    w, V, P = orc.dense_reference_evd(K)
    top = res.lam[::-1][:32]
    print("dense top", w[:4], "gpu", top[:4], flush=True)
    assert np.max(np.abs(top - w[:32])) < 1e-8


def test_sinkhorn_mode(gpu_ctx):
    X, eps, out, F = _factor(gpu_ctx, n=1200, rank=150, seed=2)
    k = 16
    res = gpu_ctx.evd(out.rank, k_out=k, prec="fp64", norm="sinkhorn", sinkhorn_tol=1e-12,
                      sinkhorn_max_iter=5000)
    lam, U, s, it = orc.lowrank_sinkhorn_evd(F, k, tol=1e-12, max_iter=5000)
    print("sinkhorn iters", res.sinkhorn_iters, it, "dev", res.rowsum_dev, flush=True)
    assert res.rowsum_dev <= 1e-12
    np.testing.assert_allclose(res.lam[-k:], lam[-k:], rtol=1e-8, atol=1e-12)
    assert abs(res.lam[-1] - 1.0) < 1e-8


@pytest.mark.parametrize("eps_scale", [0.25, 0.1])
def test_sinkhorn_small_bandwidth_schedule(gpu_ctx, eps_scale):
    """Narrow kernels (lambda_2 of P at 0.95 and 0.9995 here) push the
    Chebyshev schedule through its restarts.  The FP64 run must reach the
    damped oracle's fixed point and take the oracle schedule's pass count
    (the same coefficients from the same deviations; one pass of slack for
    a restart test decided differently by rounding)."""
    X = clustered_points(1500, 8, seed=7)
    gpu_ctx.set_points(X)
    eps = orc.median_sqdist(X) * eps_scale
    out = gpu_ctx.arpc(eps, 200, block=32, seed=7, prec="fp64")
    F = gpu_ctx.get_factor(out.rank, "fp64")
    k = 16
    res = gpu_ctx.evd(out.rank, k_out=k, prec="fp64", norm="sinkhorn", sinkhorn_tol=1e-12,
                      sinkhorn_max_iter=5000)
    lam_d, _, _, it_d = orc.lowrank_sinkhorn_evd(F, k, tol=1e-12, max_iter=20000)
    _, _, _, it_c = orc.lowrank_sinkhorn_evd(F, k, tol=1e-12, max_iter=20000, method="chebyshev")
    print("eps x", eps_scale, "gpu passes", res.sinkhorn_iters, "oracle chebyshev", it_c, "damped", it_d,
          flush=True)
    assert res.rowsum_dev <= 1e-12
    np.testing.assert_allclose(res.lam[-k:], lam_d[-k:], rtol=1e-8, atol=1e-12)
    assert abs(res.sinkhorn_iters - it_c) <= 1
    assert res.sinkhorn_iters < it_d


@pytest.mark.parametrize("n,rank", [(1200, 150), (6000, 600)])
def test_sinkhorn_mixed_vs_oracle(gpu_ctx, n, rank):
    # FP32 factor: one Gram F^T S^2 F (lower-triangle tiles, ragged edge at
    # rank 600) against the FP64 oracle on the same FP32-rounded factor
    X, eps, out, F = _factor(gpu_ctx, n=n, rank=rank, seed=4)
    F32 = F.astype(np.float32)
    gpu_ctx.set_factor(F32)
    k = 32
    res = gpu_ctx.evd(out.rank, k_out=k, prec="mixed", norm="sinkhorn", sinkhorn_tol=1e-12,
                      sinkhorn_max_iter=5000)
    lam, U, s, it = orc.lowrank_sinkhorn_evd(F32.astype(np.float64), k, tol=1e-12, max_iter=5000)
    rel = np.abs(res.lam[-k:] - lam[-k:]) / np.abs(lam[-k:])
    print("mixed sinkhorn rank", rank, "iters", res.sinkhorn_iters, it, "top-32 max rel err", rel.max(),
          flush=True)
    assert res.rowsum_dev <= 1e-12
    assert rel.max() < 1e-5
    Ua = _align(res.U.astype(np.float64), U)
    gaps = np.abs(np.diff(lam[-k - 1:]))
    ok = (gaps[:-1] > 1e-4) & (gaps[1:] > 1e-4)
    err = np.max(np.abs(Ua - U), axis=0)
    assert np.all(err[-8:][ok[-8:]] < 1e-3), err[-8:]


def test_nonpositive_degree_status(gpu_ctx):
    X = gauss_points(400, 3, seed=8)
    gpu_ctx.set_points(X)
    F = np.zeros((400, 5))
    F[:, 0] = 1.0
    F[7, 0] = -3.0  # d~_7 = F_7 . (F^T 1) < 0
    gpu_ctx.set_factor(F)
    from kernel_evd_amd import KevdError
    with pytest.raises(KevdError) as e:
        gpu_ctx.evd(5, k_out=2, prec="fp64")
    assert e.value.status == 2


def test_leading_solver_matches_full_eigensolve(gpu_ctx):
    X, eps, out, F = _factor(gpu_ctx, n=4000, dim=8, rank=700, seed=6)
    k = 100
    full = gpu_ctx.evd(out.rank, k_out=k, prec="fp64", all_eigenvalues=True)
    lead = gpu_ctx.evd(out.rank, k_out=k, prec="fp64", all_eigenvalues=False)
    assert np.all(np.isfinite(lead.lam[-k:]))
    # the full divide-and-conquer solve is accurate to ~1e-16 lambda_max in
    # absolute terms, so tiny eigenvalues are compared with that floor
    diff = np.abs(lead.lam[-k:] - full.lam[-k:])
    print("leading vs full: max abs", diff.max(), "lam_k", full.lam[-k], flush=True)
    assert np.all(diff <= 1e-9 * np.abs(full.lam[-k:]) + 1e-14)
    lam, U, _, _ = orc.lowrank_evd(F, k)
    assert np.all(np.abs(lead.lam[-k:] - lam[-k:]) <= 1e-8 * np.abs(lam[-k:]) + 1e-14)
    Ua = _align(lead.U.astype(np.float64), U)
    assert np.max(np.abs(Ua[:, -16:] - U[:, -16:])) < 1e-5


def _known_spectrum_factor(n, r, lam, seed):
    """F = [1, Q diag(sqrt(n lam))] with Q orthonormal and orthogonal to 1:
    K^ 1 = n 1, so the Sinkhorn scaling is 1/sqrt(n) and the bistochastic
    matrix P = 11^T/n + Q diag(lam) Q^T has eigenvalues 1 and lam."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, r - 1))
    G -= G.mean(axis=0)
    Q, _ = np.linalg.qr(G)
    F = np.empty((n, r))
    F[:, 0] = 1.0
    F[:, 1:] = Q * np.sqrt(lam * n)
    return F


@pytest.mark.parametrize("k", [32, 100, 200])
def test_leading_solve_known_spectrum(k):
    """Block subspace iteration on a spectrum lam_j = 0.5 * 0.97^j that it
    resolves in a few steps.  The blocks have l = 128, 256 and 288 columns:
    the first two orthonormalise with the register-resident Cholesky
    (l <= 256), the last with the panel one.  Each must converge in its own
    iteration (no Householder restart, no full-solve fallback) to the exact
    eigenvalues."""
    from kernel_evd_amd import Context
    n, r = 4096, 600
    lam = 0.5 * 0.97 ** np.arange(r - 1)
    F = _known_spectrum_factor(n, r, lam, seed=23)
    exact = np.concatenate([[1.0], lam])[:k]
    with Context(0) as c:
        c.set_points(gauss_points(n, 3, seed=24))
        c.set_factor(F)
        e = c.evd(r, k_out=k, norm="sinkhorn", prec="fp64", want_U=False)
        tm = c.timings()
    top = e.lam[::-1][:k]
    print("k", k, "iters", tm["eig_iters"], "passes", tm["eig_orth_passes"], "householder",
          tm["eig_householder"], "fallback", tm["eig_full_fallback"], "max rel err",
          np.max(np.abs(top - exact) / exact), flush=True)
    assert tm["eig_iters"] > 0 and tm["eig_orth_passes"] > 0
    assert tm["eig_householder"] == 0 and tm["eig_full_fallback"] == 0
    np.testing.assert_allclose(top, exact, rtol=1e-10)


def test_leading_solve_falls_back_on_slowly_decaying_spectrum(gpu_ctx):
    """A factor whose bistochastic spectrum is known: F = [1, Q diag(sqrt(n
    lam))] with Q orthonormal and orthogonal to 1 gives K^ 1 = n 1, so the
    Sinkhorn scaling is 1/sqrt(n) and P = 11^T/n + Q diag(lam) Q^T, with
    eigenvalues 1 and lam_j = 0.5 * 0.9997^j.  For k = 100 the block
    subspace iteration (l = 256 columns) contracts by lam_257/lam_100 ~ 0.954
    per step and cannot converge within its step limit; the full solve must
    supply the leading pairs."""
    from kernel_evd_amd import Context
    n, r, k = 4096, 600, 100
    lam = 0.5 * 0.9997 ** np.arange(r - 1)
    F = _known_spectrum_factor(n, r, lam, seed=21)
    exact = np.concatenate([[1.0], lam])[:k]
    with Context(0) as c:
        c.set_points(gauss_points(n, 3, seed=22))
        c.set_factor(F)
        e = c.evd(r, k_out=k, norm="sinkhorn", prec="fp64", want_U=False)
        tm = c.timings()
    top = e.lam[::-1][:k]
    print("fallback", tm["eig_full_fallback"], "iters", tm["eig_iters"], "max rel err",
          np.max(np.abs(top - exact) / exact), flush=True)
    assert tm["eig_full_fallback"] == 1
    np.testing.assert_allclose(top, exact, rtol=1e-10)
