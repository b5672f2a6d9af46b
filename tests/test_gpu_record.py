"""The kEVD's intermediate record (kevd_set_dump / kevd_get_intermediate):
the reference's kevd_dump datasets (ks_kevd.cu:531-542, 809-815, 935-941,
1183-1218, 1274-1280, 1562-1576, 1705-1744, 1945-1969, 2164-2170) in its
kernel basis K(:,S).

Bars, FP64 factor form against the oracle's kevd_record (the reference's
passes over K(:,S) with L^-1 from a dense Cholesky, no factor F) run in
long double: the reference's route in FP64 loses eps kappa(K(S,S))^1.3 in
R and G (2e-7 in G here, kappa = 2.8e6), the F form 400 times less.
  c1, dtil, cq, qtil          1e-10 relative (max norm)
  p1, pq                      1e-8 relative (they carry L^-T)
  Atil, A, Btil, W            1e-10 relative Frobenius
  R 1e-9, G 1e-8              relative Frobenius (G carries L^-1)
  lam                         1e-10 absolute
  V, Z, U columns with a relative eigenvalue gap above 1e-4: 1e-6 of
  their scale, up to sign.
The mixed factor records its basis as FP32 holds it: dtil, qtil and lam to
1e-5.  The LM form's panel kEVD (no factor held) computes its Grams in the
kernel basis and carries them through L^-1 as the reference does: it
records the factor form's c1, dtil, cq, qtil, Atil, Btil and lam to 1e-9,
and p1, pq, A, R, W to 64 eps kappa, G (one more L^-1) to
64 eps kappa^1.5, kappa = cond(K(S,S)).  Errors: a factor without pivots,
nothing recorded, a Sinkhorn run, too small a capacity.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import clustered_points, gauss_points

import oracle as orc
from kernel_evd_amd import KevdError
from kernel_evd_amd._lib import KEVD_EINVAL

pytestmark = pytest.mark.gpu

VEC = ("c1", "dtil", "cq", "qtil")
SOLVED = ("p1", "pq")
MATS = ("Atil", "A", "R", "G", "Btil", "W")
MAT_BAR = {"Atil": 1e-10, "A": 1e-10, "Btil": 1e-10, "W": 1e-10, "R": 1e-9, "G": 1e-8}


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _gapped(lam, k):
    """Columns (of the trailing k) whose eigenvalue is separated from both
    neighbours by more than 1e-4 relative."""
    w = lam[-k:]
    g = np.abs(np.diff(w)) / np.abs(w[-1])
    ok = np.ones(k, dtype=bool)
    ok[0] = False
    ok[:-1] &= g > 1e-4
    ok[1:] &= g > 1e-4
    return ok


def _cols_match(A, B, ok, bar):
    for j in np.flatnonzero(ok):
        a, b = A[:, j], B[:, j]
        e = min(np.abs(a - b).max(), np.abs(a + b).max()) / np.abs(b).max()
        assert e < bar, (j, e)


@pytest.fixture(scope="module")
def case():
    X = gauss_points(3000, 4, seed=31)
    eps = orc.median_sqdist(X)
    return X, eps


def test_record_fp64_matches_reference_basis(gpu_ctx, case):
    X, eps = case
    k = 12
    gpu_ctx.set_points(X)
    out = gpu_ctx.arpc(eps, 240, block=32, seed=5, prec="fp64")
    gpu_ctx.set_dump(True)
    try:
        e = gpu_ctx.evd(out.rank, k_out=k, norm="ref", prec="fp64", all_eigenvalues=True)
        rec = gpu_ctx.intermediates()
    finally:
        gpu_ctx.set_dump(False)
    ref = orc.kevd_record(X, eps, out.pivots, k, extended=True)
    r = out.rank
    for name in VEC + SOLVED:
        rel = np.abs(rec[name] - ref[name]).max() / np.abs(ref[name]).max()
        print(name, "%.2e" % rel, flush=True)
        assert rel < (1e-8 if name in SOLVED else 1e-10), name
    for name in MATS:
        assert rec[name].shape == (r, r), name
        rel = _rel(rec[name], ref[name])
        print(name, "%.2e" % rel, flush=True)
        assert rel < MAT_BAR[name], name
    assert np.all(np.triu(rec["Atil"], 1) == 0.0) and np.all(np.tril(rec["R"], -1) == 0.0)
    np.testing.assert_allclose(rec["lam"], ref["lam"], atol=1e-10)
    np.testing.assert_array_equal(rec["lam"], e.lam)
    assert rec["V"].shape == (r, r) and rec["Z"].shape == (r, k)
    ok = _gapped(ref["lam"], k)
    assert ok.sum() >= 4
    _cols_match(rec["V"][:, r - k:], ref["V"][:, r - k:], ok, 1e-6)
    _cols_match(rec["Z"], ref["Z"], ok, 1e-6)
    _cols_match(e.U.astype(np.float64), ref["U"], ok, 1e-6)


def test_record_leading_only_keeps_the_computed_block(gpu_ctx):
    X = clustered_points(20000, 6, seed=32)
    gpu_ctx.set_points(X)
    out = gpu_ctx.arpc(orc.median_sqdist(X), 600, block=64, seed=6, prec="fp64")
    gpu_ctx.set_dump(True)
    try:
        e = gpu_ctx.evd(out.rank, k_out=8, norm="ref", prec="fp64")
        V = gpu_ctx.intermediate("V")
        W = gpu_ctx.intermediate("W")
        lam = gpu_ctx.intermediate("lam")
    finally:
        gpu_ctx.set_dump(False)
    r = out.rank
    assert V.shape == (r, 8)
    np.testing.assert_array_equal(lam, e.lam)
    # W V = V diag(lam) for the computed block
    res = np.abs(W @ V - V * lam[-8:]).max()
    assert res < 1e-9 * abs(lam[-1]), res


def test_record_mixed_factor(gpu_ctx, case):
    X, eps = case
    gpu_ctx.set_points(X)
    out = gpu_ctx.arpc(eps, 240, block=32, seed=5, prec="mixed")
    gpu_ctx.set_dump(True)
    try:
        e = gpu_ctx.evd(out.rank, k_out=8, norm="ref", prec="mixed", all_eigenvalues=True)
        rec = gpu_ctx.intermediates()
    finally:
        gpu_ctx.set_dump(False)
    ref = orc.kevd_record(X, eps, out.pivots, 8)
    for name in ("dtil", "qtil"):
        rel = np.abs(rec[name] - ref[name]).max() / np.abs(ref[name]).max()
        print(name, "%.2e" % rel, flush=True)
        assert rel < 1e-5, name
    assert np.abs(rec["lam"][-8:] - ref["lam"][-8:]).max() < 1e-5
    np.testing.assert_array_equal(rec["lam"], e.lam)


def test_record_panels_match_factor_form(gpu_ctx):
    X = clustered_points(70000, 8, seed=33)
    eps = orc.median_sqdist(X, seed=0, nsample=200001)
    gpu_ctx.set_points(X)
    gpu_ctx.set_dump(True)
    try:
        lm = gpu_ctx.arpc(eps, 200, block=32, seed=7, prec="fp64", regenerate=True)
        gpu_ctx.evd(lm.rank, k_out=8, norm="ref", prec="fp64", all_eigenvalues=True)
        rec_lm = gpu_ctx.intermediates()
        ff = gpu_ctx.arpc(eps, 200, block=32, seed=7, prec="fp64", inject_pivots=lm.pivots)
        np.testing.assert_array_equal(ff.pivots, lm.pivots)
        gpu_ctx.evd(ff.rank, k_out=8, norm="ref", prec="fp64", all_eigenvalues=True)
        rec_ff = gpu_ctx.intermediates()
    finally:
        gpu_ctx.set_dump(False)
    S = lm.pivots
    kappa = np.linalg.cond(orc.kernel_block(X, S, S, eps))
    e64 = np.finfo(np.float64).eps
    bars = {name: 1e-9 for name in ("c1", "dtil", "cq", "qtil", "Atil", "Btil", "lam")}
    bars.update({name: 64 * e64 * kappa for name in ("p1", "pq", "A", "R", "W")})
    bars["G"] = 64 * e64 * kappa ** 1.5
    errs = {name: _rel(rec_lm[name], rec_ff[name]) for name in bars}
    print("kappa %.2e" % kappa, " ".join("%s %.1e" % kv for kv in errs.items()), flush=True)
    for name, bar in bars.items():
        assert errs[name] < max(bar, 1e-9), name


def test_record_errors(gpu_ctx):
    X = gauss_points(1000, 3, seed=34)
    gpu_ctx.set_points(X)
    eps = orc.median_sqdist(X)
    out = gpu_ctx.arpc(eps, 50, block=16, seed=1, prec="fp64")
    F = gpu_ctx.get_factor(out.rank, "fp64")
    gpu_ctx.evd(out.rank, k_out=4, prec="fp64")
    with pytest.raises(KevdError):  # not recording
        gpu_ctx.intermediate("A")
    gpu_ctx.set_dump(True)
    try:
        gpu_ctx.evd(out.rank, k_out=4, norm="sinkhorn", prec="fp64")
        with pytest.raises(KevdError):  # Sinkhorn records nothing
            gpu_ctx.intermediate("c1")
        gpu_ctx.evd(out.rank, k_out=4, prec="fp64")
        n = C.c_int64(3)
        buf = np.zeros(3)
        st = gpu_ctx._lib.kevd_get_intermediate(gpu_ctx._h, b"c1",
                                                buf.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n))
        assert st == KEVD_EINVAL  # capacity below r
        with pytest.raises(KevdError):
            gpu_ctx.intermediate("nosuch")
        gpu_ctx.set_factor(F)  # no pivots: no kernel basis
        with pytest.raises(KevdError) as ei:
            gpu_ctx.evd(out.rank, k_out=4, prec="fp64")
        assert ei.value.status == KEVD_EINVAL
    finally:
        gpu_ctx.set_dump(False)
