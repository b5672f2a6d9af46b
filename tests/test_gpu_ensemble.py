"""GPU tests of the batched ensemble path (kevd_ensemble_arpc / _evd) and the
block-size autotuner.

Bars: every problem of a batched mixed run gives the same pivots, accepted
counts, trace errors and factor bits as kevd_arpc on the same points and
parameters (both run the same round bodies); the FP64 ensemble matches the
CPU oracle's pivots exactly and its factor to 1e-9; the concurrent EVDs
match single-context EVDs to 1e-10 relative (the ensemble's batched
leading-eigenpair solve and a problem's own solve stop at different
steps); the batched solve's eigenvalues match the problems' own solves
(KEVD_ENS_EIG_BATCH=0) to 1e-10 relative (both stop at Ritz residuals of
1e-6 theta, so eigenvalues agree to about 1e-12) and the leading
eigenvectors with a relative gap above 1e-4 to 1e-5 of their scale, up to
sign; the autotuner returns one
of its candidates with positive measured rates.
"""
import numpy as np
import pytest

from conftest import clustered_points

import oracle as orc

pytestmark = pytest.mark.gpu

# (n, dim, rank, block): uneven sizes, dims padded to 32 / 64 / 128, ragged n
PROBLEMS = [(2000, 24, 120, 32), (5003, 64, 300, 128), (3100, 100, 200, 64), (1500, 8, 90, 16),
            (4096, 40, 256, 48)]


def _contexts(specs, seed0=0):
    from kernel_evd_amd import Context
    ctxs, Xs, eps = [], [], []
    for i, (n, dim, _, _) in enumerate(specs):
        X = clustered_points(n, dim, seed=seed0 + i).astype(np.float32)
        c = Context(0)
        c.set_points(X)
        ctxs.append(c)
        Xs.append(X)
        eps.append(orc.median_sqdist(X.astype(np.float64), seed=i, nsample=20001))
    return ctxs, Xs, eps


def test_ensemble_mixed_matches_single_problem_runs():
    from kernel_evd_amd import Context, ensemble_arpc, ensemble_evd
    ctxs, Xs, eps = _contexts(PROBLEMS)
    ranks = [p[2] for p in PROBLEMS]
    blocks = [p[3] for p in PROBLEMS]
    seeds = [3 + i for i in range(len(PROBLEMS))]
    outs, used_graph = ensemble_arpc(ctxs, eps, ranks, blocks, seeds=seeds, prec="mixed")
    print("ensemble ranks", [o.rank for o in outs], "graph", used_graph, flush=True)
    assert used_graph
    evs = ensemble_evd(ctxs, [o.rank for o in outs], k_out=12, norm="sinkhorn", prec="mixed", workers=3)
    # the Sinkhorn iterations ran batched over the problems, with each
    # problem's own schedule (compared with its single-problem run below)
    assert all(c.timings()["sinkhorn_batched"] == 1 for c in ctxs)
    for i, (X, o) in enumerate(zip(Xs, outs)):
        assert o.rank == ranks[i]
        with Context(0) as c:
            c.set_points(X)
            ref = c.arpc(eps[i], ranks[i], block=blocks[i], seed=seeds[i], prec="mixed")
            np.testing.assert_array_equal(o.pivots, ref.pivots)
            np.testing.assert_array_equal(o.accepted, ref.accepted)
            np.testing.assert_array_equal(o.trace_err, ref.trace_err)
            Fe = ctxs[i].get_factor(o.rank, "mixed")
            Fr = c.get_factor(ref.rank, "mixed")
            assert np.array_equal(Fe, Fr), "problem %d factor differs" % i
            er = c.evd(ref.rank, k_out=12, prec="mixed", norm="sinkhorn")
            assert c.timings()["sinkhorn_batched"] == 0
            assert evs[i].sinkhorn_iters == er.sinkhorn_iters
            np.testing.assert_allclose(evs[i].lam[-12:], er.lam[-12:], rtol=1e-10, atol=0)
            assert evs[i].rowsum_dev <= 1e-12
    for c in ctxs:
        c.close()


def test_ensemble_batched_eigensolve_matches_own_solves(monkeypatch):
    from kernel_evd_amd import ensemble_arpc, ensemble_evd
    ctxs, Xs, eps = _contexts(PROBLEMS, seed0=10)
    ranks = [p[2] for p in PROBLEMS]
    outs, _ = ensemble_arpc(ctxs, eps, ranks, [p[3] for p in PROBLEMS], seeds=list(range(len(PROBLEMS))),
                            prec="mixed")
    rk = [o.rank for o in outs]
    monkeypatch.setenv("KEVD_ENS_EIG_BATCH", "0")
    monkeypatch.setenv("KEVD_ENS_SK_BATCH", "0")
    own = ensemble_evd(ctxs, rk, k_out=12, norm="sinkhorn", prec="mixed", workers=3, want_U=True)
    assert all(c.timings()["sinkhorn_batched"] == 0 for c in ctxs)
    monkeypatch.delenv("KEVD_ENS_EIG_BATCH", raising=False)
    monkeypatch.delenv("KEVD_ENS_SK_BATCH", raising=False)
    bat = ensemble_evd(ctxs, rk, k_out=12, norm="sinkhorn", prec="mixed", workers=3, want_U=True)
    tms = [c.timings() for c in ctxs]
    print("ranks", rk, "eig iters", [t["eig_iters"] for t in tms], "batched", [t["eig_batched"] for t in tms],
          "full fallback", [t["eig_full_fallback"] for t in tms], flush=True)
    # k_out 12: block l = 96, so ranks >= 192 join the batch and the smaller
    # ones take the full solve; a joined problem the batch did not converge
    # by the step limit solves alone (its own solve then falls back to the
    # full one)
    for i, r in enumerate(rk):
        if r < 192:
            assert tms[i]["eig_batched"] == 0
        else:
            assert tms[i]["eig_batched"] == 1 or tms[i]["eig_full_fallback"] == 1
    assert sum(t["eig_batched"] for t in tms) >= 1
    for i in range(len(PROBLEMS)):
        np.testing.assert_allclose(bat[i].lam[-12:], own[i].lam[-12:], rtol=1e-10, atol=0)
        assert bat[i].rowsum_dev <= 1e-12
        w = own[i].lam[-12:]
        gap = np.abs(np.diff(w)) / np.abs(w[-1])
        ok = np.ones(12, dtype=bool)
        ok[0] = False
        ok[:-1] &= gap > 1e-4
        ok[1:] &= gap > 1e-4
        Ub, Uo = bat[i].U.astype(np.float64), own[i].U.astype(np.float64)
        for j in np.flatnonzero(ok):
            e = min(np.abs(Ub[:, j] - Uo[:, j]).max(), np.abs(Ub[:, j] + Uo[:, j]).max()) / np.abs(Uo[:, j]).max()
            assert e < 1e-5, (i, j, e)
    for c in ctxs:
        c.close()


def test_ensemble_fp64_matches_oracle():
    from kernel_evd_amd import ensemble_arpc
    specs = [(1800, 12, 100, 16), (2500, 20, 150, 64)]
    ctxs, Xs, eps = _contexts(specs, seed0=20)
    outs, _ = ensemble_arpc(ctxs, eps, [s[2] for s in specs], [s[3] for s in specs], seeds=[11, 12],
                            prec="fp64")
    for i, (X, o) in enumerate(zip(Xs, outs)):
        ref = orc.arpc(X.astype(np.float64), eps[i], specs[i][2], block=specs[i][3], seed=11 + i)
        assert o.rank == ref.rank
        np.testing.assert_array_equal(o.pivots, ref.pivots)
        np.testing.assert_allclose(o.trace_err, ref.trace_err, rtol=1e-9, atol=1e-14)
        F = ctxs[i].get_factor(o.rank, "fp64")
        assert np.max(np.abs(F - ref.F)) < 1e-9
    for c in ctxs:
        c.close()


def test_ensemble_rejects_mixed_precisions_and_injection():
    from kernel_evd_amd import KevdError, ensemble_arpc
    from kernel_evd_amd import _lib as L
    specs = [(1000, 8, 40, 16), (1200, 8, 40, 16)]
    ctxs, Xs, eps = _contexts(specs, seed0=40)
    import ctypes as C
    lib = L.lib()
    ps = (L.ArpcParams * 2)()
    rs = (L.ArpcResult * 2)()
    keep = []
    for i in range(2):
        ps[i].eps = eps[i]
        ps[i].rank = 40
        ps[i].block = 16
        ps[i].tol = 1e-10
        ps[i].prec = L.KEVD_MIXED if i == 0 else L.KEVD_FP64
        piv = np.zeros(40, dtype=np.int64)
        acc = np.zeros(64, dtype=np.int32)
        err = np.zeros(64)
        keep.append((piv, acc, err))
        rs[i].pivots = piv.ctypes.data_as(C.POINTER(C.c_int64))
        rs[i].accepted = acc.ctypes.data_as(C.POINTER(C.c_int32))
        rs[i].trace_err = err.ctypes.data_as(C.POINTER(C.c_double))
        rs[i].rounds_cap = 64
    hs = (C.c_void_p * 2)(*[c._h.value for c in ctxs])
    assert lib.kevd_ensemble_arpc(hs, 2, ps, rs, None) == L.KEVD_EINVAL
    assert b"same precision" in lib.kevd_last_error(ctxs[0]._h)
    ps[1].prec = L.KEVD_MIXED
    inj = np.arange(10, dtype=np.int64)
    ps[1].inject_pivots = inj.ctypes.data_as(C.POINTER(C.c_int64))
    ps[1].n_inject = 10
    assert lib.kevd_ensemble_arpc(hs, 2, ps, rs, None) == L.KEVD_EINVAL
    with pytest.raises(KevdError):
        ensemble_arpc([ctxs[0], ctxs[0]], eps, [40, 40], [16, 16])
    for c in ctxs:
        c.close()


def test_autotune_block():
    from kernel_evd_amd import Context
    X = clustered_points(6000, 32, seed=9).astype(np.float32)
    with Context(0) as c:
        c.set_points(X)
        eps = c.median_sqdist(seed=0, nsample=20001)
        best, rates = c.autotune_block(eps, 300, candidates=(16, 64, 128), measure=True)
        print("autotune", best, rates, flush=True)
        assert best in (16, 64, 128)
        assert np.all(rates > 0)
        assert best == (16, 64, 128)[int(np.argmax(rates))]
        assert c.autotune_block(eps, 300, candidates=(16, 64, 128)) == best  # cached


def test_autotune_tiles_chooses_and_keeps_the_results():
    """Both Gram tiles are measured on the factor and the faster one is
    kept; the kEVD's eigenvalues do not depend on the tile beyond FP32
    accumulation order (1e-6 relative)."""
    from kernel_evd_amd import Context, KevdError
    X = clustered_points(20000, 16, seed=40).astype(np.float32)
    eps = orc.median_sqdist(X.astype(np.float64), seed=0, nsample=20001)
    with Context(0) as c:
        c.set_points(X)
        a = c.arpc(eps, 300, block=64, seed=1, prec="mixed")
        tile, ms = c.autotune_tiles(measure=True)
        print("tile", tile, "ms (256, 128)", ms, flush=True)
        assert tile in (128, 256) and np.all(ms > 0)
        assert tile == (128 if ms[1] < ms[0] else 256)
        assert c.autotune_tiles() == tile  # cached for this shape
        lams = {}
        for t in (256, 128):
            c.set_gram_tile(t)
            lams[t] = c.evd(a.rank, k_out=16, prec="mixed").lam[-16:]
        np.testing.assert_allclose(lams[128], lams[256], rtol=1e-6)
        with pytest.raises(KevdError):
            c.set_gram_tile(64)
    with Context(0) as c:
        c.set_points(X)
        c.arpc(eps, 300, block=64, seed=1, prec="fp64")
        assert c.autotune_tiles() == 0  # FP64 factor: no choice
