"""Pins the CPU oracle (oracle/kevd_oracle.c, oracle/oracle.py) against
known answers, textbook identities and brute force -- never against a
retyped copy of its own formulas.

  Philox4x32-10       Random123 known-answer vectors (Salmon et al. SC'11)
  kernel block        scipy cdist brute force; K_ii = 1; symmetry
  median              numpy.median over scipy pdist (exact mode)
  sampler             one-hot residual; zero mass; chi-square of frequencies
  rejection block     H = I accepts all; duplicates rejected; numpy cholesky
  ARPC                injected pivots: F = K(:,S) chol(K(S,S))^-T (textbook);
                      full rank: F F^T = K; interpolation; trace bookkeeping
  bandwidth sums      brute force; eps -> inf and eps -> 0 limits
  kEVD (NumPy)        full rank equals the dense ks_reference restatement;
                      lambda_max = 1, P 1 = 1 (bistochastic by construction)
  kEVD record         the reference-basis passes (kevd_record): d~, q~ are
                      the Nystrom degrees K(:,S) K(S,S)^-1 K(S,:) 1 by a
                      direct solve; p1, pq solve K(S,S); R^T R = A + sI;
                      lam and U equal the F-form restatement's
  Sinkhorn            s .* (K s) = 1 at convergence
  KS ETDRK4           linear-regime closed form exp(L t); mean conservation;
                      4th-order self-convergence in dt
"""
import numpy as np
import pytest
from scipy.spatial.distance import cdist, pdist

from conftest import clustered_points, gauss_points

import oracle as orc


def test_philox_known_answers():
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    m = 0xffffffff
    assert orc.philox([m, m, m, m], [m, m]) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                      [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_range():
    u = [orc.uniform01(7, orc.STREAM_REJECT, 3, i) for i in range(2000)]
    assert min(u) > 0.0 and max(u) <= 1.0
    assert abs(np.mean(u) - 0.5) < 0.03


def test_kernel_block_brute_force():
    X = gauss_points(120, 7, seed=1)
    eps = 5.0
    rows = np.arange(0, 120, 3)
    cols = np.arange(120)
    K = orc.kernel_block(X, rows, cols, eps)
    ref = np.exp(-cdist(X[rows], X, "sqeuclidean") / eps)
    assert np.max(np.abs(K - ref)) < 1e-15
    for i, r in enumerate(rows):
        assert K[i, r] == 1.0
    Kf = orc.kernel_block(X, cols, cols, eps)
    assert np.array_equal(Kf, Kf.T)


def test_median_exact_and_sampled():
    X = gauss_points(301, 5, seed=2)
    assert orc.median_sqdist(X) == np.median(pdist(X, "sqeuclidean"))
    X2 = gauss_points(300, 5, seed=2)  # even pair count
    assert abs(orc.median_sqdist(X2) - np.median(pdist(X2, "sqeuclidean"))) < 1e-12
    a = orc.median_sqdist(X, seed=1, nsample=5001)
    b = orc.median_sqdist(X, seed=2, nsample=5001)
    full = np.median(pdist(X, "sqeuclidean"))
    assert abs(a - full) < 0.1 * full and a != b


def test_sampler_known_answers():
    d = np.zeros(1000)
    d[417] = 0.3
    S, W = orc.propose(d, seed=5, rnd=0, b=64)
    assert np.all(S == 417) and W > 0
    S, W = orc.propose(np.zeros(50), seed=5, rnd=0, b=4)
    assert S is None and W == 0
    # weights and prefix: W < 2^62, prefix non-decreasing
    W, pre = orc.fixed_weights(np.random.default_rng(0).random(5000))
    assert W < 2 ** 62 and np.all(np.diff(pre.astype(np.float64)) >= 0)


def test_sampler_distribution():
    d = np.array([0.0, 1.0, 2.0, 0.0, 3.0, 4.0])
    counts = np.zeros(6)
    for rnd in range(400):
        S, _ = orc.propose(d, seed=9, rnd=rnd, b=32)
        np.add.at(counts, S, 1)
    assert counts[0] == 0 and counts[3] == 0
    p = d / d.sum()
    exp = p * counts.sum()
    chi2 = np.sum((counts[p > 0] - exp[p > 0]) ** 2 / exp[p > 0])
    assert chi2 < 16.3  # 3 dof, p = 0.001


def test_rejchol_identity_accepts_all():
    k, acc, Lc, Li = orc.rejchol(np.eye(16), seed=1, rnd=0)
    assert k == 16 and np.array_equal(acc, np.arange(16))
    assert np.allclose(Lc, np.eye(16)) and np.allclose(Li, np.eye(16))


def test_rejchol_duplicates_and_cholesky():
    X = gauss_points(40, 3, seed=3)
    idx = np.array([0, 1, 2, 1, 3, 4, 5, 6, 2, 7])
    H = orc.kernel_block(X, idx, idx, 2.0)
    k, acc, Lc, Li = orc.rejchol(H.copy(), seed=2, rnd=0, tol=1e-10, inject=True)
    assert 3 not in acc and 8 not in acc           # repeated proposals: residual 0
    Hs = H[np.ix_(acc, acc)]
    assert np.allclose(Lc, np.linalg.cholesky(Hs), atol=1e-13)
    assert np.allclose(Li @ Lc, np.eye(k), atol=1e-12)


def test_arpc_injected_pivots_textbook():
    X = clustered_points(500, 6, seed=4)
    eps = orc.median_sqdist(X)
    S = np.random.default_rng(1).choice(500, 37, replace=False)
    r = orc.arpc(X, eps, 37, block=8, inject=S)
    assert np.array_equal(r.pivots, S)
    KSS = orc.kernel_block(X, S, S, eps)
    KnS = orc.kernel_block(X, np.arange(500), S, eps)
    L = np.linalg.cholesky(KSS)
    F = np.linalg.solve(L, KnS.T).T                 # K(:,S) L^-T
    assert np.max(np.abs(r.F - F)) < 1e-10


def test_arpc_full_rank_and_bookkeeping():
    X = gauss_points(80, 3, seed=6)
    eps = 3.0 * orc.median_sqdist(X)
    r = orc.arpc(X, eps, 80, block=16, seed=1, tol=0.0)
    K = orc.dense_kernel(X, eps)
    assert np.max(np.abs(r.F @ r.F.T - K)) < 1e-8 * 80 + 10 * r.trace_err[-1]
    X = clustered_points(1500, 8, seed=7)
    eps = orc.median_sqdist(X)
    r = orc.arpc(X, eps, 120, block=32, seed=3)
    KS = orc.kernel_block(X, np.arange(1500), r.pivots, eps)
    assert np.max(np.abs(r.F @ r.F[r.pivots].T - KS)) < 1e-12
    assert np.allclose(r.d, np.maximum(1.0 - np.sum(r.F ** 2, axis=1), 0.0), atol=1e-12)
    assert abs(r.trace_err[-1] - r.d.mean()) < 1e-14
    assert np.all(np.diff(r.trace_err) <= 1e-15)
    assert r.accepted.sum() == r.rank == 120
    # b = 1 accepts every proposal (u * rnd < u unless rnd == 1)
    r1 = orc.arpc(X, eps, 30, block=1, seed=3)
    assert r1.rounds == 30 and np.all(r1.accepted == 1)


def test_bandwidth_sums_brute_force_and_limits():
    X = gauss_points(200, 4, seed=8)
    rows = np.arange(0, 200, 7)
    eps = np.array([1e-6, 0.5, 2.0, 1e9])
    S = orc.bandwidth(X, rows, eps)
    D = cdist(X[rows], X, "sqeuclidean")
    ref = np.array([np.exp(-D / e).sum() for e in eps])
    assert np.allclose(S, ref, rtol=1e-13)
    assert abs(S[0] - rows.size) < 1e-9              # only self pairs survive
    assert abs(S[-1] - rows.size * 200) < 1e-3       # every pair -> 1


def test_lowrank_evd_full_rank_equals_dense():
    X = gauss_points(150, 3, seed=9)
    eps = 2.0 * orc.median_sqdist(X)
    r = orc.arpc(X, eps, 150, block=16, seed=2, tol=0.0)
    lam, U, dev, s = orc.lowrank_evd(r.F, 20)
    w, V, P = orc.dense_reference_evd(orc.dense_kernel(X, eps))
    assert np.max(np.abs(lam[::-1][:20] - w[:20])) < 1e-10
    assert abs(w[0] - 1.0) < 1e-12 and np.max(np.abs(P.sum(axis=1) - 1.0)) < 1e-12
    assert np.max(np.abs(dev)) < 1e-10
    # eigenvectors agree up to sign
    for j in range(1, 6):
        u, v = U[:, -1 - j], V[:, j]
        assert min(np.abs(u - v).max(), np.abs(u + v).max()) < 1e-6 * np.abs(v).max() * 150


def test_kevd_record_reference_basis():
    X = gauss_points(300, 3, seed=21)
    eps = orc.median_sqdist(X)
    ref = orc.arpc(X, eps, 60, block=16, seed=1)
    S, k = ref.pivots, 8
    rec = orc.kevd_record(X, eps, S, k)
    K = orc.kernel_block(X, np.arange(300), S, eps)
    KSS = K[S, :]
    # Nystrom degrees by a direct solve, no Cholesky factor involved
    Kn = K @ np.linalg.solve(KSS, K.T)
    np.testing.assert_allclose(rec["dtil"], Kn @ np.ones(300), rtol=1e-9)
    np.testing.assert_allclose(rec["qtil"], Kn @ (1.0 / rec["dtil"]), rtol=1e-9)
    np.testing.assert_allclose(KSS @ rec["p1"], rec["c1"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(KSS @ rec["pq"], rec["cq"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(rec["c1"], K.sum(axis=0), rtol=1e-12)
    # Atil: lower triangle of K^T D^-2 K; R upper with R^T R = A + sI
    assert np.all(np.triu(rec["Atil"], 1) == 0.0) and np.all(np.tril(rec["R"], -1) == 0.0)
    Kd = K / rec["dtil"][:, None]
    np.testing.assert_allclose(rec["Atil"], np.tril(Kd.T @ Kd), rtol=1e-9, atol=1e-12)
    r = S.size
    np.testing.assert_allclose(rec["R"].T @ rec["R"], rec["A"] + rec["shift"] * np.eye(r), rtol=1e-9, atol=1e-12)
    # the same spectrum and eigenvectors as the F-form restatement on the C
    # oracle's factor F = K(:,S) L^-T
    lam, U, dev, sF = orc.lowrank_evd(ref.F, k)
    np.testing.assert_allclose(rec["lam"], lam, atol=1e-11)
    assert abs(rec["lam"][-1] - 1.0) < 1e-10 and abs(rec["shift"] - sF) < 1e-12 * abs(sF) + 1e-30
    for j in range(k):
        u, v = rec["U"][:, j], U[:, j]
        assert min(np.abs(u - v).max(), np.abs(u + v).max()) < 1e-7 * np.abs(v).max()


def test_kevd_record_extended_precision():
    # rank 240 of 3000 points: kappa(K(S,S)) = 2.8e6.  The FP64 F-form
    # route (F from the C oracle, R = chol(F^T D~^-2 F + sI), G L = R by a
    # solve with L = F(S,:)) and the long double kernel-basis route agree
    # in G to 5e-10; the FP64 kernel-basis route is 400x further off.
    X = gauss_points(3000, 4, seed=31)
    eps = orc.median_sqdist(X)
    ref = orc.arpc(X, eps, 240, block=32, seed=5)
    S, F = ref.pivots, ref.F
    ext = orc.kevd_record(X, eps, S, 8, extended=True)
    dbl = orc.kevd_record(X, eps, S, 8)
    n, r = F.shape
    dt = F @ (F.T @ np.ones(n))
    Af = (F / dt[:, None]).T @ (F / dt[:, None])
    Rf = np.linalg.cholesky(Af + np.finfo(float).eps * np.trace(Af) * np.eye(r)).T
    Gf = np.linalg.solve(F[S, :].T, Rf.T).T
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(Gf, ext["G"]) < 5e-9
    assert rel(Rf, ext["R"]) < 1e-10
    assert rel(dbl["G"], ext["G"]) > 1e-8
    for name in ("c1", "dtil", "cq", "qtil", "Atil", "Btil", "W", "lam"):
        assert rel(dbl[name], ext[name]) < 1e-10, name


def test_sinkhorn_fixed_point():
    X = clustered_points(400, 5, seed=10)
    eps = orc.median_sqdist(X)
    r = orc.arpc(X, eps, 80, block=16, seed=4)
    lam, U, s, it = orc.lowrank_sinkhorn_evd(r.F, 8, tol=1e-13)
    Ks = r.F @ (r.F.T @ s)
    assert np.max(np.abs(s * Ks - 1.0)) <= 1e-13
    assert abs(lam[-1] - 1.0) < 1e-10


@pytest.mark.parametrize("eps_scale", [1.0, 0.25])
def test_sinkhorn_chebyshev_schedule_reaches_the_damped_fixed_point(eps_scale):
    """The accelerated schedule (the library's) and the damped iteration
    converge to the same scaling; the fixed point s .* (K^ s) = 1 is unique,
    so any error in the schedule's coefficients either stalls it or lands
    elsewhere, and both would fail here."""
    X = clustered_points(1500, 6, seed=11)
    eps = orc.median_sqdist(X) * eps_scale
    r = orc.arpc(X, eps, 150, block=32, seed=5)
    lam_d, _, s_d, it_d = orc.lowrank_sinkhorn_evd(r.F, 8, tol=1e-12, max_iter=5000)
    lam_c, _, s_c, it_c = orc.lowrank_sinkhorn_evd(r.F, 8, tol=1e-12, max_iter=5000, method="chebyshev")
    Ks = r.F @ (r.F.T @ s_c)
    assert np.max(np.abs(s_c * Ks - 1.0)) <= 1e-12
    assert np.max(np.abs(s_c - s_d) / s_d) < 1e-10
    assert np.max(np.abs(lam_c[-8:] - lam_d[-8:])) < 1e-10
    print("damped", it_d, "chebyshev", it_c)
    assert it_c < it_d


def _ks_lambda(L, M):
    q = 2 * np.pi / L * np.arange(M // 2 + 1)
    return q ** 2 - q ** 4


def test_ks_linear_regime_closed_form():
    M, L, dt, steps = 64, 22.0, 0.25, 40
    x = np.arange(M)
    u0 = 1e-9 * np.cos(2 * np.pi * 3 * x / M)       # single mode k = 3
    out = orc.ks_numpy(2, L=L, M=M, dt=dt, n_discard=0, stride=steps, u0=u0)
    lam = _ks_lambda(L, M)[3]
    exact = np.exp(lam * dt * steps) * u0
    assert np.max(np.abs(out[1] - exact)) < 1e-6 * np.max(np.abs(exact))


def test_ks_mean_conserved_and_fourth_order():
    M, L = 64, 22.0
    u0 = orc.gauss_ic(M, 0)
    u0 = u0 + 0.3                                    # nonzero mean
    out = orc.ks_numpy(3, L=L, M=M, dt=0.25, n_discard=0, stride=40, u0=u0)
    assert np.allclose(out.mean(axis=1), u0.mean(), atol=1e-12)
    # self-convergence from the smooth Kassam-Trefethen initial condition
    # (white noise shows the order reduction ETDRK4 has on rough data)
    x = L * np.arange(M) / M
    us = np.cos(2 * np.pi * x / L) * (1 + np.sin(2 * np.pi * x / L))
    T = 4.0
    sols = [orc.ks_numpy(2, L=L, M=M, dt=dt, n_discard=0, stride=int(round(T / dt)), u0=us)[1]
            for dt in (0.2, 0.1, 0.05, 0.025)]
    e = [np.max(np.abs(s - sols[-1])) for s in sols[:-1]]
    # against the finest run: 16 for 4th order asymptotically, ~8 for 3rd
    assert 7.0 < e[0] / e[1] < 20.0 and 7.0 < e[1] / e[2] < 20.0


def test_gauss_ic_is_standard_normal():
    u = np.concatenate([orc.gauss_ic(64, s, scale=1.0) for s in range(60)])
    assert abs(u.mean()) < 0.08 and abs(u.std() - 1.0) < 0.06


def test_dense_reference_leading_subset_equals_full():
    X = clustered_points(300, 4, seed=13)
    K = orc.dense_kernel(X, orc.median_sqdist(X))
    w, _, _ = orc.dense_reference_evd(K)
    wk, Vk, _ = orc.dense_reference_evd(K, k=12)
    assert np.max(np.abs(wk - w[:12])) < 1e-13 and abs(wk[0] - 1.0) < 1e-12
    assert Vk.shape == (300, 12)


def test_arpc_stops_when_no_residual_exceeds_tol():
    X = clustered_points(500, 3, seed=14)
    eps = orc.median_sqdist(X)
    # tol above every residual entry (all start at K_ii = 1): nothing to accept
    r = orc.arpc(X, eps, 50, block=16, seed=1, tol=2.0)
    assert r.rank == 0 and r.rounds == 0
    # five distinct points, forty copies each: once one copy of each is a
    # pivot every residual entry is at rounding level (below tol) and the
    # run stops at rank 5 instead of running out its round budget
    base = gauss_points(5, 4, seed=15)
    Xd = np.repeat(base, 40, axis=0)
    r = orc.arpc(Xd, orc.median_sqdist(base), 50, block=16, seed=2)
    assert r.rank == 5 and r.rounds < 10
    assert len(set(map(tuple, Xd[r.pivots]))) == 5
    assert np.all(np.isfinite(r.F)) and r.trace_err[-1] < 1e-12


def _post(eps, S):
    from kernel_evd_amd import bandwidth_post
    return bandwidth_post(eps, S)


def test_bandwidth_post_closed_forms():
    """ks_bandwidth.cu:609-642 on sums with a known slope.  S = c eps^(m/2):
    every centred slope is m/2, so mdim = m (the location of the maximum is
    decided by rounding there).
    log S = (sqrt(pi)/2) erf(log eps - x0): the slope peaks at x0, where the
    centred difference over a grid step h is sqrt(pi) erf(h) / (2 h)."""
    eps = np.logspace(-2, 2, 41)
    logS, slope, eopt, mdim = _post(eps, 3.0 * eps ** 2.5)
    np.testing.assert_allclose(slope, 2.5, rtol=1e-12)
    assert abs(mdim - 5.0) < 1e-11
    np.testing.assert_allclose(logS, np.log(3.0) + 2.5 * np.log(eps), rtol=1e-13)
    from scipy.special import erf
    x = np.log(eps)
    x0 = x[23]
    h = x[1] - x[0]
    logS, slope, eopt, mdim = _post(eps, np.exp(0.5 * np.sqrt(np.pi) * erf(x - x0)))
    assert eopt == eps[23]
    assert abs(mdim - np.sqrt(np.pi) * erf(h) / h) < 1e-10
