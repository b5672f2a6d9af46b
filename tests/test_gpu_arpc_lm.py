"""GPU tests of the LM form of ARPC (kevd_arpc_params.regenerate = 1): the
reference's own ks_arpclm.cu algorithm, which keeps S and L^-1 and
regenerates K(R,S) every round instead of storing F.

Bars:
- FP64 LM against the CPU oracle (F-form, same Philox stream): pivots and
  accepted counts bit-exact, trace errors 1e-9, L^-1 1e-8 relative;
- FP64 LM against the GPU F-form run: the same, 1e-10 on the trace errors;
- every quantity on the L^-1 path is FP64 in both precisions, so
  L^-1 K(S,S) L^-T = I to 1e-8 for the mixed run's own pivots as well, and
  with the FP64 pivots injected the mixed LM L^-1 equals the FP64 one to
  1e-9; the mixed trace error is within 5% of FP64's;
- no factor is left behind (get_factor and evd raise).
"""
import numpy as np
import pytest

from conftest import clustered_points

import oracle as orc

pytestmark = pytest.mark.gpu


def _setup(n=3000, dim=16, seed=5):
    from kernel_evd_amd import Context
    X = clustered_points(n, dim, seed=seed)
    c = Context(0)
    c.set_points(X)
    eps = orc.median_sqdist(X, seed=0, nsample=20001)
    return c, X, eps


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("block", [1, 32, 128])
def test_lm_fp64_matches_oracle(block):
    c, X, eps = _setup()
    rank = 200
    print("lm fp64 block", block, flush=True)
    lm = c.arpc(eps, rank, block=block, seed=7, prec="fp64", regenerate=True)
    ref = orc.arpc(X, eps, rank, block=block, seed=7)
    assert lm.rank == ref.rank == rank
    np.testing.assert_array_equal(lm.pivots, ref.pivots)
    np.testing.assert_array_equal(lm.accepted, ref.accepted)
    np.testing.assert_allclose(lm.trace_err, ref.trace_err, rtol=1e-9, atol=1e-14)
    Li = c.get_linv(rank)
    Lref = np.linalg.inv(ref.F[ref.pivots, :])  # F(S,:) = L
    assert _rel(Li, Lref) < 1e-8
    c.close()


def test_lm_fp64_matches_factor_form():
    c, X, eps = _setup(n=5000, dim=24, seed=6)
    rank = 300
    ff = c.arpc(eps, rank, block=64, seed=3, prec="fp64")
    Lf = c.get_linv(rank)
    lm = c.arpc(eps, rank, block=64, seed=3, prec="fp64", regenerate=True)
    Ll = c.get_linv(rank)
    np.testing.assert_array_equal(lm.pivots, ff.pivots)
    np.testing.assert_array_equal(lm.accepted, ff.accepted)
    np.testing.assert_allclose(lm.trace_err, ff.trace_err, rtol=1e-10, atol=1e-15)
    assert _rel(Ll, Lf) < 1e-9
    c.close()


def test_lm_mixed_linv_exact_and_trace_error():
    from kernel_evd_amd import KevdError
    c, X, eps = _setup(n=6000, dim=32, seed=8)
    rank = 400
    r64 = c.arpc(eps, rank, block=128, seed=1, prec="fp64", regenerate=True)
    L64 = c.get_linv(rank)
    mx = c.arpc(eps, rank, block=128, seed=1, prec="mixed", regenerate=True)
    assert mx.rank == rank
    Lm = c.get_linv(rank)
    K = c.kernel_block(mx.pivots, mx.pivots, eps, prec="fp64")
    E = Lm @ K @ Lm.T - np.eye(rank)
    print("mixed LM: |L K L^T - I| max", np.abs(E).max(), "trace err", mx.trace_err[-1], r64.trace_err[-1],
          flush=True)
    assert np.abs(E).max() < 1e-8
    assert abs(mx.trace_err[-1] - r64.trace_err[-1]) <= 0.05 * r64.trace_err[-1]
    # the FP64 pivots injected: the L^-1 path is all FP64
    inj = c.arpc(eps, rank, block=128, seed=1, prec="mixed", regenerate=True, inject_pivots=r64.pivots)
    np.testing.assert_array_equal(inj.pivots, r64.pivots)
    assert _rel(c.get_linv(rank), L64) < 1e-9
    # no factor kept
    with pytest.raises(KevdError):
        c.get_factor(rank, "mixed")
    with pytest.raises(KevdError):
        c.evd(rank, k_out=8, prec="mixed")
    # a factor-form run afterwards works as usual
    ff = c.arpc(eps, rank, block=128, seed=1, prec="mixed")
    assert c.get_factor(ff.rank, "mixed").shape == (6000, rank)
    c.close()


def test_lm_rejected_by_ensemble():
    from kernel_evd_amd import Context
    from kernel_evd_amd import _lib as L
    import ctypes as C
    X = clustered_points(1000, 8, seed=2)
    with Context(0) as c0, Context(0) as c1:
        c0.set_points(X)
        c1.set_points(X)
        lib = L.lib()
        ps = (L.ArpcParams * 2)()
        rs = (L.ArpcResult * 2)()
        keep = []
        for i in range(2):
            ps[i].eps = 10.0
            ps[i].rank = 20
            ps[i].block = 8
            ps[i].tol = 1e-10
            ps[i].prec = L.KEVD_MIXED
            ps[i].regenerate = i
            piv = np.zeros(20, dtype=np.int64)
            acc = np.zeros(64, dtype=np.int32)
            err = np.zeros(64)
            keep.append((piv, acc, err))
            rs[i].pivots = piv.ctypes.data_as(C.POINTER(C.c_int64))
            rs[i].accepted = acc.ctypes.data_as(C.POINTER(C.c_int32))
            rs[i].trace_err = err.ctypes.data_as(C.POINTER(C.c_double))
            rs[i].rounds_cap = 64
        hs = (C.c_void_p * 2)(c0._h.value, c1._h.value)
        assert lib.kevd_ensemble_arpc(hs, 2, ps, rs, None) == L.KEVD_EINVAL
        assert b"LM form" in lib.kevd_last_error(c0._h)
