"""GPU tests on the five benchmark configurations at their full sizes.

  C2 + C3  N = 1,048,576 KS snapshots (dim 64, FP32), eps = the median squared
           distance, rank 2000, block 128, seed 0, mixed beside FP64; then the
           bistochastic kEVD (Sinkhorn to 1e-12 and the two-step form), top 100
  C1       16,384 seeded rows x all N = 262,144 snapshots x 64 eps
           log-spaced over [1e-3, 1e3] x median, one fused launch
  C0       N = 8192 snapshots (FP64), exact median, block 1, rank 512; and the
           same run to the rank where no residual entry exceeds tol = 1e-10
  C4       48 KS problems, N ~ U[4096, 131072], (L, M) in {(22, 32), (44, 64),
           (88, 128)}, rank ~ U[256, 1024] (seed 1), per-problem median eps,
           autotuned block size, batched

The CPU oracle checks what it can finish in seconds: the pivots and factor
on prefixes (8192 and 65,536 points), the sweep on 256 rows, the dense
two-step EVD at N = 8192, and the smallest ensemble problems.  At full size
the checks are size-independent properties: interpolation at the pivots,
the trace bookkeeping, monotone trace error, lambda_max = 1, the row-sum
deviation, and agreement of the mixed and FP64 paths (ks_arpclm.cu:793-1046,
ks_kevd.cu:755-2308, ks_bandwidth.cu:181-325, ks_reference.py:21-40).
"""
import time

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

_KS = {}


def ks(n):
    """The first n snapshots of the KS trajectory (L = 22, 64 points,
    dt = 0.25, IC 0.1 N(0,1) seed 0, 10,000 steps discarded, stride 4).  The
    generator is sequential in time, so a prefix of a longer run is the
    shorter run and one generation serves every size."""
    from kernel_evd_amd import ks_generate
    for m, X in _KS.items():
        if m >= n:
            return X[:n]
    t = time.time()
    X = ks_generate(n, L_dom=22.0, M=64, dt=0.25, n_discard=10000, stride=4, seed=0)
    print("generated %d KS snapshots in %.1f s" % (n, time.time() - t), flush=True)
    _KS.clear()
    _KS[n] = X
    return X


def _rel_top(a, b, k):
    a, b = a[::-1][:k], b[::-1][:k]
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.timeout(900)
def test_c2_c3_rank2000_factor_and_evd():
    from kernel_evd_amd import Context
    n, rank, block = 1 << 20, 2000, 128
    X32 = np.ascontiguousarray(ks(n), dtype=np.float32)
    with Context(0) as c, Context(0) as c64:
        c.set_points(X32)
        c64.set_points(X32)
        eps = c.median_sqdist(seed=0, nsample=1 << 22)
        print("C2: eps", eps, flush=True)
        a = c.arpc(eps, rank, block=block, seed=0, prec="mixed")
        a64 = c64.arpc(eps, rank, block=block, seed=0, prec="fp64")
        print("C2: mixed rank %d in %d rounds, trace err %.4e; fp64 %.4e"
              % (a.rank, a.rounds, a.trace_err[-1], a64.trace_err[-1]), flush=True)
        assert a.rank == rank and a64.rank == rank
        for t in (a.trace_err, a64.trace_err):
            assert np.all(np.diff(t) <= 0.0) and t[-1] > 0.0
        # the reference algorithm's trace error, within 5% (the two runs take
        # different pivot sequences)
        assert abs(a.trace_err[-1] - a64.trace_err[-1]) <= 0.05 * a64.trace_err[-1]
        # C3: Sinkhorn to 1e-12, top 100, and the two-step form
        e = c.evd(rank, k_out=100, norm="sinkhorn", prec="mixed", want_U=False)
        e64 = c64.evd(rank, k_out=100, norm="sinkhorn", prec="fp64", want_U=False)
        print("C3: sinkhorn iters %d, row-sum dev %.2e, top-32 rel err %.2e, top-100 %.2e"
              % (e.sinkhorn_iters, e.rowsum_dev, _rel_top(e.lam, e64.lam, 32), _rel_top(e.lam, e64.lam, 100)),
              flush=True)
        assert e.rowsum_dev <= 1e-12 and e64.rowsum_dev <= 1e-12
        # lambda_max = 1: to FP64 rounding on the FP64 path, to the FP32
        # Gram's rounding on the mixed one
        assert abs(e.lam[-1] - 1.0) < 1e-7 and abs(e64.lam[-1] - 1.0) < 1e-12
        assert _rel_top(e.lam, e64.lam, 32) < 1e-5
        er = c.evd(rank, k_out=100, norm="ref", prec="mixed", want_U=False)
        er64 = c64.evd(rank, k_out=100, norm="ref", prec="fp64", want_U=False)
        print("C3: two-step top-32 rel err %.2e, row-sum dev %.2e" % (_rel_top(er.lam, er64.lam, 32), er.rowsum_dev),
              flush=True)
        assert abs(er64.lam[-1] - 1.0) < 1e-10 and er64.rowsum_dev < 1e-9
        assert _rel_top(er.lam, er64.lam, 32) < 1e-5
        # factor properties at full size
        rows = np.sort(np.random.default_rng(5).choice(n, 2048, replace=False))
        for ctx, out, prec, bar in ((c, a, "mixed", 1e-4), (c64, a64, "fp64", 1e-10)):
            F = ctx.get_factor(rank, prec)
            fro2 = 0.0
            for i in range(0, n, 1 << 16):
                blk = F[i:i + (1 << 16)].astype(np.float64)
                fro2 += float(np.sum(blk * blk))
            # tr(K - F F^T) / tr(K) = 1 - |F|_F^2 / N, with d >= 0 clamping
            # (which only adds) as the one difference
            te = out.trace_err[-1]
            gap = te - (1.0 - fro2 / n)
            print("C2 %s: trace err %.6e, 1 - |F|^2/N gap %.2e" % (prec, te, gap), flush=True)
            assert -1e-12 <= gap <= 1e-3 * te
            # interpolation: F F(S,:)^T = K(:,S)
            KS = ctx.kernel_block(rows, out.pivots, eps, "fp64")
            G = F[rows].astype(np.float64) @ F[out.pivots].astype(np.float64).T
            ierr = float(np.abs(G - KS).max())
            print("C2 %s: interpolation max |F F(S)^T - K(:,S)| = %.2e" % (prec, ierr), flush=True)
            assert ierr < bar
            del F, G, KS


@pytest.mark.timeout(600)
def test_c2_fp64_pivots_match_oracle_on_prefix():
    """FP64 pivots bit for bit against the oracle on the first 65,536
    snapshots at rank 2000, block 128 (the full-size oracle run would take
    minutes)."""
    from kernel_evd_amd import Context
    X = ks(65536)
    eps = orc.median_sqdist(X, seed=0, nsample=1 << 20)
    t = time.time()
    ref = orc.arpc(X, eps, 2000, block=128, seed=0)
    print("oracle ARPC 65536 x 2000 in %.1f s, trace err %.3e" % (time.time() - t, ref.trace_err[-1]), flush=True)
    with Context(0) as c:
        c.set_points(X)
        out = c.arpc(eps, 2000, block=128, seed=0, prec="fp64")
        assert out.rank == ref.rank == 2000
        np.testing.assert_array_equal(out.pivots, ref.pivots)
        np.testing.assert_array_equal(out.accepted, ref.accepted)
        np.testing.assert_allclose(out.trace_err, ref.trace_err, rtol=1e-9, atol=1e-15)
        F = c.get_factor(out.rank, "fp64")
        assert np.max(np.abs(F - ref.F)) < 1e-8


@pytest.mark.timeout(600)
def test_c1_bandwidth_sweep_full():
    from kernel_evd_amd import Context, bandwidth_post
    n, nrows, neps = 262144, 16384, 64
    X = np.ascontiguousarray(ks(n), dtype=np.float32)
    rows = np.sort(np.random.default_rng(0).choice(n, nrows, replace=False)).astype(np.int64)
    with Context(0) as c:
        c.set_points(X)
        med = c.median_sqdist(seed=0, nsample=1 << 22)
        eps = med * np.logspace(-3, 3, neps)
        S32 = c.bandwidth_sweep(rows, eps, "mixed")
        S64 = c.bandwidth_sweep(rows, eps, "fp64")
        sub = rows[:256]
        Ssub32 = c.bandwidth_sweep(sub, eps, "mixed")
        Ssub64 = c.bandwidth_sweep(sub, eps, "fp64")
        lim = c.bandwidth_sweep(rows, np.array([1e-9 * med, 1e9 * med]), "mixed")
    rel = float(np.max(np.abs(S32 - S64) / S64))
    ref = orc.bandwidth(X.astype(np.float64), sub, eps)
    print("C1: mixed vs fp64 max rel %.2e; 256-row oracle: mixed %.2e, fp64 %.2e"
          % (rel, np.max(np.abs(Ssub32 - ref) / ref), np.max(np.abs(Ssub64 - ref) / ref)), flush=True)
    assert rel < 2e-6
    np.testing.assert_allclose(Ssub64, ref, rtol=1e-11)
    assert np.max(np.abs(Ssub32 - ref) / ref) < 2e-6
    _, sl32, eopt32, mdim32 = bandwidth_post(eps, S32)
    _, sl64, eopt64, mdim64 = bandwidth_post(eps, S64)
    assert np.max(np.abs(sl32 - sl64)) < 1e-5
    assert eopt32 == eopt64 and abs(mdim32 - mdim64) < 1e-5
    # eps -> 0 keeps the self pairs (distinct snapshots), eps -> inf every pair
    assert abs(lim[0] - nrows) < 1e-6 * nrows
    assert abs(lim[1] - float(nrows) * n) < 1e-5 * nrows * n


@pytest.mark.timeout(300)
def test_c0_small_fp64_block1_rank512():
    from kernel_evd_amd import Context
    X = ks(8192)
    eps = orc.median_sqdist(X)  # exact over all pairs
    ref = orc.arpc(X, eps, 512, block=1, seed=0)
    with Context(0) as c:
        c.set_points(X)
        assert abs(c.median_sqdist() - eps) <= 1e-12 * eps
        out = c.arpc(eps, 512, block=1, seed=0, prec="fp64")
        print("C0: rank %d in %d rounds, trace err %.3e" % (out.rank, out.rounds, out.trace_err[-1]), flush=True)
        assert out.rank == 512 and out.rounds == 512
        np.testing.assert_array_equal(out.pivots, ref.pivots)
        np.testing.assert_allclose(out.trace_err, ref.trace_err, rtol=1e-9, atol=1e-15)
        F = c.get_factor(512, "fp64")
        assert np.max(np.abs(F - ref.F)) < 1e-9


@pytest.mark.timeout(600)
def test_c0_to_tolerance_matches_dense_reference():
    """Block-1 FP64 run continued until no residual entry exceeds tol =
    1e-10 (trace error <= 1e-10); its two-step top-32 eigenvalues against the
    dense FP64 restatement of ks_reference.py:21-40."""
    from kernel_evd_amd import Context
    n = 8192
    X = ks(n)
    eps = orc.median_sqdist(X)
    with Context(0) as c:
        c.set_points(X)
        out = c.arpc(eps, n, block=1, seed=0, tol=1e-10, prec="fp64", max_rounds=n + 64)
        print("C0: stopped at rank %d after %d rounds, trace err %.3e"
              % (out.rank, out.rounds, out.trace_err[-1]), flush=True)
        assert out.rank < n and out.trace_err[-1] <= 1e-10
        e = c.evd(out.rank, k_out=32, norm="ref", prec="fp64", want_U=False)
    K = orc.dense_kernel(X, eps)
    t = time.time()
    w = orc.dense_reference_evd(K, k=32)[0]
    print("C0: dense reference top-32 in %.1f s" % (time.time() - t), flush=True)
    rel = np.abs(e.lam[::-1][:32] - w[:32]) / np.abs(w[:32])
    print("C0: top-32 rel err vs dense %.2e" % rel.max(), flush=True)
    assert rel.max() < 1e-8


ENS_SHAPES = [(22.0, 32), (44.0, 64), (88.0, 128)]


@pytest.mark.timeout(900)
def test_c4_ensemble_48_problems():
    from concurrent.futures import ThreadPoolExecutor
    from kernel_evd_amd import Context, ensemble_arpc, ensemble_evd, ks_generate
    nprob = 48
    rng = np.random.default_rng(1)
    Ns = rng.integers(4096, 131072 + 1, nprob)
    shp = rng.integers(0, len(ENS_SHAPES), nprob)
    ranks = [int(v) for v in rng.integers(256, 1024 + 1, nprob)]
    t = time.time()
    with ThreadPoolExecutor(16) as ex:
        data = list(ex.map(lambda i: ks_generate(int(Ns[i]), L_dom=ENS_SHAPES[shp[i]][0], M=ENS_SHAPES[shp[i]][1],
                                                 dt=0.25, n_discard=10000, stride=4, seed=1000 + i)
                           .astype(np.float32), range(nprob)))
    print("C4: generated %d data sets (%d points) in %.1f s" % (nprob, int(Ns.sum()), time.time() - t), flush=True)
    ctxs = [Context(0) for _ in range(nprob)]
    try:
        eps, blocks = [], []
        for i, c in enumerate(ctxs):
            c.set_points(data[i])
            eps.append(c.median_sqdist(seed=i, nsample=1 << 20))
            blocks.append(c.autotune_block(eps[i], ranks[i], candidates=(32, 64, 128), seed=i))
        seeds = list(range(nprob))
        print("C4: blocks chosen", {b: blocks.count(b) for b in sorted(set(blocks))}, flush=True)
        outs, graph = ensemble_arpc(ctxs, eps, ranks, blocks, seeds=seeds, prec="mixed")
        assert graph
        evs = ensemble_evd(ctxs, [o.rank for o in outs], k_out=32, norm="sinkhorn", prec="mixed")
        with Context(0) as s:
            for i in range(nprob):
                s.set_points(data[i])
                ref = s.arpc(eps[i], ranks[i], block=blocks[i], seed=seeds[i], prec="mixed")
                o = outs[i]
                assert o.rank == ranks[i] == ref.rank, i
                np.testing.assert_array_equal(o.pivots, ref.pivots)
                np.testing.assert_array_equal(o.accepted, ref.accepted)
                np.testing.assert_array_equal(o.trace_err, ref.trace_err)
                assert np.array_equal(ctxs[i].get_factor(o.rank, "mixed"), s.get_factor(ref.rank, "mixed")), i
                er = s.evd(ref.rank, k_out=32, norm="sinkhorn", prec="mixed", want_U=False)
                # the two solves may stop at different steps (the first Rayleigh-Ritz
                # check follows each context's last solve): ~1e-12 relative
                np.testing.assert_allclose(evs[i].lam[-32:], er.lam[-32:], rtol=1e-10, atol=0)
        # FP64 ensemble: the three smallest problems against the oracle, and
        # every problem's top-32 against its mixed run on the same pivots
        o64, _ = ensemble_arpc(ctxs, eps, ranks, blocks, seeds=seeds, prec="fp64")
        for i in np.argsort(Ns)[:3]:
            X64 = data[i].astype(np.float64)
            r = orc.arpc(X64, eps[i], ranks[i], block=blocks[i], seed=seeds[i])
            np.testing.assert_array_equal(o64[i].pivots, r.pivots)
            assert np.max(np.abs(ctxs[i].get_factor(o64[i].rank, "fp64") - r.F)) < 1e-9
        errs = []
        for i, c in enumerate(ctxs):
            c.arpc(eps[i], outs[i].rank, block=blocks[i], seed=seeds[i], prec="fp64", inject_pivots=outs[i].pivots)
            e64 = c.evd(outs[i].rank, k_out=32, norm="sinkhorn", prec="fp64", want_U=False)
            errs.append(_rel_top(evs[i].lam, e64.lam, 32))
        print("C4: top-32 rel err vs FP64 on the same pivots: median %.2e, max %.2e"
              % (np.median(errs), max(errs)), flush=True)
        assert max(errs) < 1e-4
    finally:
        for c in ctxs:
            c.close()
