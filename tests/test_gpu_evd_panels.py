"""The kEVD with no factor held (the reference's LM form, ks_kevd.cu:755-800,
1091-1168, 1627-1690, 2229-2275): after kevd_arpc with regenerate = 1 the
context keeps only the pivots and L^-1, and kevd_evd regenerates K(R,S) in
row panels of 65,536 points for every pass, in FP64.  Bars: eigenvalues
within 1e-9 relative of the factor-form kEVD on the same pivots (both
normalisations), its leading eigenvectors to 1e-5 of their scale, the
row-sum deviation as the factor form's; an FP32 request is refused."""
import numpy as np
import pytest

from conftest import clustered_points

import oracle as orc

pytestmark = pytest.mark.gpu


def _align(U, V):
    s = np.sign(np.sum(U * V, axis=0))
    s[s == 0] = 1
    return U * s


@pytest.fixture(scope="module")
def lm_case():
    # three panels, the last one partial
    X = clustered_points(150000, 10, seed=41).astype(np.float32)
    eps = orc.median_sqdist(X.astype(np.float64), seed=0, nsample=200001)
    return X, eps


@pytest.mark.parametrize("norm", ["ref", "sinkhorn"])
def test_panel_evd_fp64_matches_factor_form(lm_case, norm):
    from kernel_evd_amd import Context
    X, eps = lm_case
    rank, k = 300, 16
    with Context(0) as c:
        c.set_points(X)
        lm = c.arpc(eps, rank, block=64, seed=4, prec="fp64", regenerate=True)
        e_lm = c.evd(lm.rank, k_out=k, norm=norm, prec="fp64")
        ff = c.arpc(eps, rank, block=64, seed=4, prec="fp64", inject_pivots=lm.pivots)
        np.testing.assert_array_equal(ff.pivots, lm.pivots)
        e_ff = c.evd(ff.rank, k_out=k, norm=norm, prec="fp64")
    rel = np.abs(e_lm.lam[-k:] - e_ff.lam[-k:]) / np.abs(e_ff.lam[-k:])
    print(norm, "panel vs factor form: max rel eig err %.2e, row-sum dev %.1e / %.1e"
          % (rel.max(), e_lm.rowsum_dev, e_ff.rowsum_dev), flush=True)
    assert rel.max() < 1e-9
    assert abs(e_lm.lam[-1] - 1.0) < 1e-9
    Ua = _align(e_lm.U.astype(np.float64), e_ff.U.astype(np.float64))
    # eigenvectors with a relative gap to both neighbours (the lowest
    # computed one has no computed neighbour below)
    gaps = np.abs(np.diff(e_ff.lam[-k:]))
    ok = np.ones(k, dtype=bool)
    ok[0] = False
    ok[:-1] &= gaps > 1e-4
    ok[1:] &= gaps > 1e-4
    err = np.max(np.abs(Ua - e_ff.U), axis=0)
    assert np.all(err[ok] < 1e-5 * np.max(np.abs(e_ff.U))), err


def test_panel_evd_is_fp64_only(lm_case):
    """In the kernel's basis the rounding of K is carried through L^-1
    (norm up to tol^-1/2): the factor-free kEVD refuses FP32 and runs in
    FP64, also after a mixed LM run."""
    from kernel_evd_amd import Context, KevdError
    from kernel_evd_amd._lib import KEVD_EINVAL
    X, eps = lm_case
    with Context(0) as c:
        c.set_points(X)
        lm = c.arpc(eps, 200, block=64, seed=5, prec="mixed", regenerate=True)
        with pytest.raises(KevdError) as ei:
            c.evd(lm.rank, k_out=8, norm="sinkhorn", prec="mixed", want_U=False)
        assert ei.value.status == KEVD_EINVAL
        e = c.evd(lm.rank, k_out=8, norm="sinkhorn", prec="fp64", want_U=False)
        assert e.rowsum_dev <= 1e-12 and abs(e.lam[-1] - 1.0) < 1e-9


def test_panel_evd_needs_the_lm_run():
    from kernel_evd_amd import Context, KevdError
    X = clustered_points(2000, 4, seed=42)
    with Context(0) as c:
        c.set_points(X)
        with pytest.raises(KevdError):
            c.evd(10, k_out=4, prec="fp64")  # nothing computed yet
