"""Failure handling on the GPU: the reference's abort paths and the edge cases
of its round loop, each turned into a status (and a driver exit code)
instead of an abort.

  status 3  Cholesky of the shifted Gram fails (ks_kevd.cu:1505-1520)
  status 4  a non-positive q~ (ks_kevd.cu:1692-1703)
  zero mass duplicate points: the residual vanishes (ks_arpclm.cu:213-223,
            where the reference's float CDF divides by a zero sum)
  tol       no residual entry above tol: the run stops, nothing accepted
  eps       NaN or non-positive bandwidths are rejected (EINVAL)

Every condition is made by the inputs; each kernel still runs to its normal
exit.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import clustered_points, gauss_points

import oracle as orc
from kernel_evd_amd import KevdError
from kernel_evd_amd._lib import KEVD_EINVAL, KEVD_ENONPOS_Q, KEVD_EPOTRF

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kernel_evd_amd", "bin")


def near_duplicate_factor(n=4096, pairs=32, seed=3):
    """Positive factor whose columns come in pairs c, c (1 + 1e-6 g): the
    weighted Gram A = F^T D~^-2 F has 32 eigenvalues ~1e-13 of its scale.
    The FP64 Gram resolves them (A + sI stays positive definite with the
    reference's shift s = eps_mach tr(A)); FP32 accumulation errors of ~1e-7
    make each pair's Schur pivot negative with probability 1/2, so the
    Cholesky of the mixed Gram fails."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(0.5, 1.5, (n, pairs))
    part = base * (1.0 + 1e-6 * rng.standard_normal((n, pairs)))
    F = np.empty((n, 2 * pairs))
    F[:, 0::2] = base
    F[:, 1::2] = part
    return F.astype(np.float32)


def overflow_factor(n=4096):
    """One column of 1e200: d~ = F F^T 1 overflows to +inf (positive), so
    d~^-1 = 0 and q~ = F F^T d~^-1 = 0: the non-positive q~ abort."""
    F = np.zeros((n, 4))
    F[:, 0] = 1e200
    return F


def test_status3_potrf_failure_on_mixed_gram(gpu_ctx):
    gpu_ctx.set_points(gauss_points(4096, 4, seed=1))
    F = near_duplicate_factor()
    gpu_ctx.set_factor(F)
    with pytest.raises(KevdError) as ei:
        gpu_ctx.evd(F.shape[1], k_out=4, norm="ref", prec="mixed", want_U=False)
    assert ei.value.status == KEVD_EPOTRF, ei.value
    # the same factor in FP64: A + sI is positive definite
    gpu_ctx.set_factor(F.astype(np.float64))
    e = gpu_ctx.evd(F.shape[1], k_out=4, norm="ref", prec="fp64", want_U=False)
    assert np.all(np.isfinite(e.lam[-4:])) and abs(e.lam[-1] - 1.0) < 1e-8


def test_status4_nonpositive_q(gpu_ctx):
    gpu_ctx.set_points(gauss_points(4096, 4, seed=2))
    gpu_ctx.set_factor(overflow_factor())
    with pytest.raises(KevdError) as ei:
        gpu_ctx.evd(4, k_out=2, norm="ref", prec="fp64", want_U=False)
    assert ei.value.status == KEVD_ENONPOS_Q, ei.value


def _run(*args):
    return subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)


def test_driver_exit_codes_3_and_4(tmp_path):
    X = tmp_path / "X.npy"
    np.save(X, gauss_points(4096, 4, seed=4))
    F3, F4 = tmp_path / "F3.npy", tmp_path / "F4.npy"
    np.save(F3, near_duplicate_factor())
    np.save(F4, overflow_factor())
    r = _run("kevd-evd", "--points", X, "--factor", F3, "--kout", 4, "--precision", "mixed")
    print(r.stdout)
    assert r.returncode == 3
    r = _run("kevd-evd", "--points", X, "--factor", F4, "--kout", 2, "--precision", "fp64")
    print(r.stdout)
    assert r.returncode == 4


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_duplicate_points_stop_cleanly(gpu_ctx, prec):
    base = gauss_points(5, 4, seed=15)
    X = np.repeat(base, 40, axis=0)
    gpu_ctx.set_points(X)
    eps = orc.median_sqdist(base)
    out = gpu_ctx.arpc(eps, 50, block=16, seed=2, prec=prec)
    F = gpu_ctx.get_factor(out.rank, "fp64")
    print(prec, "rank", out.rank, "rounds", out.rounds, "trace err", out.trace_err[-1], flush=True)
    assert np.all(np.isfinite(F)) and np.all(np.isfinite(out.trace_err))
    assert len(set(map(tuple, X[out.pivots[:5]]))) == 5
    if prec == "fp64":
        # every residual entry falls to rounding level (below tol) after the
        # fifth pivot: the run stops there, as the oracle's does
        ref = orc.arpc(X, eps, 50, block=16, seed=2)
        assert out.rank == ref.rank == 5
        np.testing.assert_array_equal(out.pivots, ref.pivots)
        np.testing.assert_array_equal(out.accepted, ref.accepted)
        assert out.rounds < 10 and out.trace_err[-1] < 1e-12
    else:
        # FP32 rounding leaves residuals of ~1e-7, above tol: the run goes on
        # to the target rank through those, with no NaN
        assert out.trace_err[-1] < 1e-5


def test_zero_tolerance_runs_to_zero_mass(gpu_ctx):
    base = gauss_points(6, 3, seed=16)
    X = np.repeat(base, 30, axis=0)
    gpu_ctx.set_points(X)
    out = gpu_ctx.arpc(orc.median_sqdist(base), 60, block=8, seed=3, tol=0.0, prec="fp64", max_rounds=200)
    F = gpu_ctx.get_factor(out.rank, "fp64")
    print("rank", out.rank, "rounds", out.rounds, "trace err", out.trace_err[-1], flush=True)
    # rounding-level residuals may be accepted (tol = 0) or rejected (a
    # computed H_ii <= 0) until the mass is exactly zero or the budget ends:
    # either way no NaN
    assert np.all(np.isfinite(F)) and out.rank >= 6
    assert np.all(np.isfinite(out.trace_err)) and out.trace_err[-1] < 1e-14


def test_tol_above_every_residual_accepts_nothing(gpu_ctx):
    X = clustered_points(1000, 5, seed=17)
    gpu_ctx.set_points(X)
    out = gpu_ctx.arpc(orc.median_sqdist(X), 40, block=16, seed=1, tol=2.0, prec="mixed")
    assert out.rank == 0 and out.rounds == 0
    with pytest.raises(KevdError):
        gpu_ctx.evd(40, k_out=4, prec="mixed")


@pytest.mark.parametrize("bad", [float("nan"), 0.0, -1.0])
def test_bad_eps_is_einval(gpu_ctx, bad):
    X = gauss_points(300, 3, seed=18)
    gpu_ctx.set_points(X)
    for call in (lambda: gpu_ctx.arpc(bad, 10, block=8, prec="fp64"),
                 lambda: gpu_ctx.kernel_block(np.arange(5), np.arange(5), bad, "mixed"),
                 lambda: gpu_ctx.bandwidth_sweep(np.arange(5), np.array([1.0, bad]), "mixed")):
        with pytest.raises(KevdError) as ei:
            call()
        assert ei.value.status == KEVD_EINVAL
