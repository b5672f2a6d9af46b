"""CPU oracle for libkevd -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product never imports it.

Two parts:
  * liboracle.so (kevd_oracle.c): FP64 Philox, kernel block, median, the
    fixed-point sampler, the rejection Cholesky and ARPC in F-form, and
    bandwidth sums -- restating ks_arpclm.cu / ks_bandwidth.cu.
  * NumPy restatements: the dense bistochastic EVD of ks_reference.py:21-40
    (``dense_reference_evd``), its F-form low-rank counterpart of
    ks_kevd.cu / README_kevd.md:22-69 (``lowrank_evd``), dense symmetric
    Sinkhorn, and the KS ETDRK4 integrator of ks_datagen.py:5-89
    (``ks_numpy``).

Pinning of each function is listed in tests/test_oracle.py.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")

_so = None
i32, i64, u32, u64, dbl = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
P = C.POINTER


def so():
    global _so
    if _so is None:
        if not os.path.exists(SO):
            import subprocess
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(SO)
        L.orc_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        L.orc_draw_u64.restype = u64
        L.orc_draw_u64.argtypes = [u64, u32, u32, u32]
        L.orc_uniform01.restype = dbl
        L.orc_uniform01.argtypes = [u64, u32, u32, u32]
        L.orc_kernel_block.argtypes = [P(dbl), i64, i32, P(i64), i64, P(i64), i64, dbl, P(dbl)]
        L.orc_median_sqdist.restype = C.c_int
        L.orc_median_sqdist.argtypes = [P(dbl), i64, i32, u64, i64, P(dbl)]
        L.orc_fixed_weights.restype = u64
        L.orc_fixed_weights.argtypes = [P(dbl), i64, P(u64)]
        L.orc_propose.restype = C.c_int
        L.orc_propose.argtypes = [P(dbl), i64, u64, u32, i32, P(i64), P(u64)]
        L.orc_rejchol.restype = i32
        L.orc_rejchol.argtypes = [P(dbl), i32, u64, u32, dbl, C.c_int, P(i32), P(dbl), P(dbl)]
        L.orc_arpc.restype = C.c_int
        L.orc_arpc.argtypes = [P(dbl), i64, i32, dbl, i64, i32, u64, dbl, i32, P(i64), i64,
                               P(dbl), P(i64), P(i32), P(dbl), P(i32), P(i64), P(dbl), P(dbl)]
        L.orc_bandwidth.argtypes = [P(dbl), i64, i32, P(i64), i64, P(dbl), i32, P(dbl)]
        L.orc_num_threads.restype = C.c_int
        _so = L
    return _so


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


STREAM_PROPOSE, STREAM_REJECT, STREAM_GAUSS, STREAM_MEDIAN, STREAM_ROWS = range(5)


def philox(ctr, key):
    c = (u32 * 4)(*ctr)
    k = (u32 * 2)(*key)
    o = (u32 * 4)()
    so().orc_philox4x32_10(c, k, o)
    return list(o)


def draw_u64(seed, stream, rnd, idx):
    return so().orc_draw_u64(seed, stream, rnd, idx)


def uniform01(seed, stream, rnd, idx):
    return so().orc_uniform01(seed, stream, rnd, idx)


def kernel_block(X, rows, cols, eps):
    X = np.ascontiguousarray(X, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty((rows.size, cols.size))
    so().orc_kernel_block(_p(X, dbl), X.shape[0], X.shape[1], _p(rows, i64), rows.size,
                          _p(cols, i64), cols.size, eps, _p(out, dbl))
    return out


def median_sqdist(X, seed=0, nsample=0):
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = C.c_double()
    rc = so().orc_median_sqdist(_p(X, dbl), X.shape[0], X.shape[1], seed, nsample, C.byref(out))
    if rc:
        raise RuntimeError("orc_median_sqdist failed")
    return out.value


def fixed_weights(d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    pre = np.empty(d.size, dtype=np.uint64)
    W = so().orc_fixed_weights(_p(d, dbl), d.size, _p(pre, u64))
    return W, pre


def propose(d, seed, rnd, b):
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.empty(b, dtype=np.int64)
    W = C.c_uint64()
    rc = so().orc_propose(_p(d, dbl), d.size, seed, rnd, b, _p(out, i64), C.byref(W))
    return (out if rc == 0 else None), W.value


def rejchol(H, seed, rnd, tol=1e-10, inject=False):
    H = np.array(H, dtype=np.float64, order="C")
    b = H.shape[0]
    acc = np.zeros(b, dtype=np.int32)
    Lc = np.zeros((b, b))
    Li = np.zeros((b, b))
    k = so().orc_rejchol(_p(H, dbl), b, seed, rnd, tol, int(inject), _p(acc, i32), _p(Lc, dbl),
                         _p(Li, dbl))
    return k, acc[:k].copy(), Lc[:k, :k].copy(), Li[:k, :k].copy()


class ArpcRef:
    pass


def arpc(X, eps, rank, block=128, seed=0, tol=1e-10, max_rounds=0, inject=None):
    """F-form ARPC in FP64 on the CPU (orc_arpc)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, dim = X.shape
    mr = max_rounds if max_rounds > 0 else 4 * ((rank + block - 1) // block) + 64
    F = np.zeros((n, rank))
    piv = np.zeros(rank, dtype=np.int64)
    acc = np.zeros(mr, dtype=np.int32)
    err = np.zeros(mr)
    rounds = C.c_int32()
    rk = C.c_int64()
    d = np.zeros(n)
    injp = None
    ninj = 0
    if inject is not None:
        inject = np.ascontiguousarray(inject, dtype=np.int64)
        injp = _p(inject, i64)
        ninj = inject.size
    phase = np.zeros(3)
    rc = so().orc_arpc(_p(X, dbl), n, dim, eps, rank, block, seed, tol, mr, injp, ninj,
                       _p(F, dbl), _p(piv, i64), _p(acc, i32), _p(err, dbl), C.byref(rounds),
                       C.byref(rk), _p(d, dbl), _p(phase, dbl))
    if rc:
        raise RuntimeError("orc_arpc failed: %d" % rc)
    r = ArpcRef()
    r.rank = rk.value
    r.rounds = rounds.value
    r.F = F[:, :r.rank]
    r.pivots = piv[:r.rank]
    r.accepted = acc[:r.rounds]
    r.trace_err = err[:r.rounds]
    r.d = d
    # wall seconds: sampling, pivot block (H + rejection), n-row downdate
    r.t_sample, r.t_block, r.t_downdate = (float(v) for v in phase)
    return r


def bandwidth(X, rows, eps):
    X = np.ascontiguousarray(X, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    S = np.zeros(eps.size)
    so().orc_bandwidth(_p(X, dbl), X.shape[0], X.shape[1], _p(rows, i64), rows.size,
                       _p(eps, dbl), eps.size, _p(S, dbl))
    return S


def num_threads():
    return so().orc_num_threads()


# ------------------------------------------------------------ dense EVD

def dense_kernel(X, eps):
    """compute_kernel_mat, ks_reference.py:51-58 (eps = Nq * bw): direct
    squared differences, FP64."""
    from scipy.spatial.distance import cdist
    X = np.asarray(X, dtype=np.float64)
    return np.exp(-cdist(X, X, "sqeuclidean") / eps)


def dense_reference_evd(K, k=None):
    """ks_reference.py:21-40: d = K1/N, q = K D^-1 1 / N, Kt = D^-1 K Q^-1/2,
    P = Kt Kt^T / N^2; eigenpairs descending.  With k, only the leading k
    pairs (LAPACK's subset solver; the same eigenvalues, less time)."""
    NM = K.shape[0]
    one = np.ones(NM)
    dvec = (1.0 / NM) * (K @ one)
    qvec = (1.0 / NM) * (K @ (one / dvec))
    Kt = (K / dvec[:, None]) / np.sqrt(qvec)[None, :]
    P = (1.0 / NM ** 2) * (Kt @ Kt.T)
    if k is None:
        w, V = np.linalg.eigh(P)
    else:
        from scipy.linalg import eigh
        w, V = eigh(P, subset_by_index=[NM - k, NM - 1], driver="evr")
    return w[::-1].copy(), V[:, ::-1].copy(), P


def lowrank_evd(F, k_out):
    """F-form of the streaming kEVD (README_kevd.md:22-69; ks_kevd.cu passes
    1-4 and dense phases 1-2), all in FP64:
      d~ = F (F^T 1); q~ = F (F^T d~^-1); A = F^T D~^-2 F; B = F^T Q~^-1 F;
      s = eps_mach tr(A); R = chol(A + sI) (upper, A = R^T R);
      W = R B R^T; W = V diag(lam) V^T; U = D~^-1 F R^-1 V_k.
    Returns lam ascending (all r), U (n x k_out, columns ascending), P1 - 1."""
    F = np.asarray(F, dtype=np.float64)
    n, r = F.shape
    dt = F @ (F.T @ np.ones(n))
    qt = F @ (F.T @ (1.0 / dt))
    Y = F / dt[:, None]
    A = Y.T @ Y
    B = F.T @ (F / qt[:, None])
    s = np.finfo(np.float64).eps * np.trace(A)
    Rl = np.linalg.cholesky(A + s * np.eye(r))  # lower, A = Rl Rl^T
    R = Rl.T
    W = R @ B @ R.T
    W = 0.5 * (W + W.T)
    lam, V = np.linalg.eigh(W)
    Vk = V[:, r - k_out:]
    Yk = np.linalg.solve(R, Vk)
    U = (F @ Yk) / dt[:, None]
    # row sums of P = Y B Y^T
    P1 = Y @ (B @ (Y.T @ np.ones(n)))
    return lam, U, P1 - 1.0, s


def _chol_lower(M):
    """Cholesky factor, lower, by columns (any float dtype, numpy has no
    long double LAPACK)."""
    n = M.shape[0]
    L = np.zeros_like(M)
    for j in range(n):
        d = M[j, j] - np.dot(L[j, :j], L[j, :j])
        if not d > 0:
            raise np.linalg.LinAlgError("not positive definite at %d" % j)
        L[j, j] = np.sqrt(d)
        L[j + 1:, j] = (M[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return L


def _inv_lower(L):
    """Inverse of a lower triangular matrix, row by row."""
    n = L.shape[0]
    X = np.zeros_like(L)
    for i in range(n):
        X[i, i] = 1 / L[i, i]
        X[i, :i] = -(L[i, :i] @ X[:i, :i]) / L[i, i]
    return X


def kevd_record(X, eps, S, k_out, extended=False):
    """The reference kEVD's dump datasets, computed as the reference does,
    over the kernel columns K = K(:,S) with L^-1 from a dense Cholesky of
    K(S,S) = L L^T (ks_kevd.cu; the dumped names of 531-542 and the later
    h5add_double sites):
      c1 = K^T 1 (755), p1 = L^-T (L^-1 c1) (883), dtil = K p1,
      cq = K^T dtil^-1 (1122), pq = L^-T (L^-1 cq) (1220), qtil = K pq,
      Atil = K^T D^-2 K (lower triangle kept, 1405), A = L^-1 Atil L^-T
      (1413-1440), s = DBL_EPSILON tr(A) (1483), R = chol(A + sI) upper
      (1500), G = R L^-1 (1538), Btil = K^T Q^-1 K (1662), W = G Btil G^T
      (1764), W = V diag(lam) V^T, Z = L^-T R^-1 V_k (1973), and
      U = D^-1 K Z (2173).
    Nothing here goes through the factor F = K L^-T.  Matrices are
    returned as 2-D FP64 arrays; Atil with its strict upper triangle zero,
    R with its strict lower triangle zero.

    This route multiplies rounding errors by L^-1 twice: in FP64, R and G
    carry relative errors up to ~eps kappa(K(S,S))^1.3 (2e-7 in G at
    kappa = 2.8e6).  extended=True runs everything up to W in long double
    (x87, eps 1.1e-19; kernel values from direct differences), which brings
    those errors below the FP64 F-form's; the eigensolve of W stays FP64."""
    S = np.asarray(S, dtype=np.int64)
    r, n = S.size, X.shape[0]
    if extended:
        T = np.longdouble
        XL = np.asarray(X, dtype=T)
        d2 = ((XL[:, None, :] - XL[None, S, :]) ** 2).sum(-1)
        K = np.exp(-d2 / T(eps))
        Lc = _chol_lower(K[S, :])
        Li = _inv_lower(Lc)
        chol = _chol_lower
    else:
        T = np.float64
        X = np.asarray(X, dtype=np.float64)
        K = kernel_block(X, np.arange(n), S, eps)
        Lc = np.linalg.cholesky(K[S, :])
        Li = np.tril(np.linalg.inv(Lc))
        chol = np.linalg.cholesky
    out = {}
    c1 = K.T @ np.ones(n, dtype=T)
    p1 = Li.T @ (Li @ c1)
    dt = K @ p1
    cq = K.T @ (1 / dt)
    pq = Li.T @ (Li @ cq)
    qt = K @ pq
    Kd = K / dt[:, None]
    Atil = Kd.T @ Kd
    A = Li @ Atil @ Li.T
    s = T(np.finfo(np.float64).eps) * np.trace(A)
    R = chol(A + s * np.eye(r, dtype=T)).T
    G = R @ Li
    Btil = K.T @ (K / qt[:, None])
    W = G @ Btil @ G.T
    W = (W + W.T) / 2
    for name, v in (("c1", c1), ("p1", p1), ("dtil", dt), ("cq", cq), ("pq", pq), ("qtil", qt),
                    ("Atil", np.tril(Atil)), ("A", A), ("R", R), ("G", G), ("Btil", Btil), ("W", W)):
        out[name] = np.asarray(v, dtype=np.float64)
    lam, V = np.linalg.eigh(out["W"])
    out["V"], out["lam"] = V, lam
    Vk = V[:, r - k_out:].astype(T)
    Z = Li.T @ np.linalg.solve(R.astype(np.float64), V[:, r - k_out:]) if not extended else \
        Li.T @ _inv_lower(R.T).T @ Vk
    out["Z"] = np.asarray(Z, dtype=np.float64)
    out["U"] = np.asarray((K @ Z) / dt[:, None], dtype=np.float64)
    out["shift"] = float(s)
    return out


def _cheb_schedule():
    """Step coefficients of the accelerated Sinkhorn iteration (the same
    schedule the library runs): after the pass that measured s_j with RMS
    deviation err, s_{j+1} = a1 s_j + a2 s_{j-1} + a3 T(s_j), where
    T(s) = sqrt(mean(s * K^ s)) / (K^ s).  Two plain steps, then the
    Chebyshev semi-iteration for I + P with the spectrum of P taken in
    [0, lam] (lam = 0.3 at first; raised and restarted when the deviation
    falls more slowly than predicted by a factor 1.5 over two steps)."""
    st = {"lam": 0.3, "j": 0, "k": 0, "rho": 0.0, "e2": []}

    def nxt(err):
        st["e2"].append(err)
        st["j"] += 1
        if st["j"] <= 2:
            return 0.0, 0.0, 1.0
        lam = st["lam"]
        th, de = 1.0 + 0.5 * lam, 0.5 * lam
        sig = th / de
        e2 = st["e2"]
        if st["k"] >= 3 and e2[-3] > 0.0:
            g = sig + np.sqrt(sig * sig - 1.0)
            q = np.sqrt(e2[-1] / e2[-3])
            if q > 1.5 / g:
                z = q * g
                x = 0.5 * (z + 1.0 / z)
                lam = min(0.995, max(1.25 * lam, 1.1 * lam * (x + 1.0) * 0.5))
                st["lam"], st["k"] = lam, 0
                th, de = 1.0 + 0.5 * lam, 0.5 * lam
                sig = th / de
        if st["k"] == 0:
            st["rho"] = 1.0 / sig
            alpha, beta = 0.0, 1.0 / th
        else:
            rn = 1.0 / (2.0 * sig - st["rho"])
            alpha, beta = rn * st["rho"], 2.0 * rn / de
            st["rho"] = rn
        st["k"] += 1
        return 1.0 - beta + alpha, -alpha, beta
    return nxt


def lowrank_sinkhorn_evd(F, k_out, tol=1e-12, max_iter=10000, method="damped"):
    """Symmetric Sinkhorn on K^ = F F^T: s .* (K^ s) = 1, P = S K^ S; the
    nonzero spectrum is that of F^T S^2 F.  method "damped": the fixed point
    iteration s <- sqrt(s / (K^ s)); "chebyshev": the accelerated schedule
    of _cheb_schedule (same fixed point, fewer passes)."""
    F = np.asarray(F, dtype=np.float64)
    n, r = F.shape
    s = np.ones(n) / np.sqrt(max(1e-300, np.mean(F @ (F.T @ np.ones(n)))))
    sprev = s
    step = _cheb_schedule() if method == "chebyshev" else None
    # the schedule falls back to the damped update after three rises of the
    # deviation in a row, as the library does
    damped, rises, prev = step is None, 0, np.inf
    it = 0
    for it in range(1, max_iter + 1):
        Ks = F @ (F.T @ s)
        rs = s * Ks
        dev = np.max(np.abs(rs - 1.0))
        if dev <= tol:
            break
        was_damped = damped
        rises = rises + 1 if dev > prev else 0
        prev = dev
        if was_damped:
            snew = np.sqrt(s / Ks)
        else:
            a1, a2, a3 = step(np.sqrt(np.mean((rs - 1.0) ** 2)))
            snew = a1 * s + a2 * sprev + a3 * np.sqrt(np.mean(rs)) / Ks
        damped = was_damped or rises >= 3
        sprev, s = s, snew
    G = F.T @ (F * (s * s)[:, None])
    G = 0.5 * (G + G.T)
    lam, V = np.linalg.eigh(G)
    Vk = V[:, r - k_out:]
    U = (s[:, None] * F) @ Vk / np.sqrt(np.maximum(lam[r - k_out:], 1e-300))[None, :]
    return lam, U, s, it


# -------------------------------------------------------------- KS data

def ks_numpy(n, L=22.0, M=64, dt=0.25, n_discard=10000, stride=4, u0=None, nsteps_total=None):
    """ETDRK4 with 3/2 dealiasing and 16-point contour coefficients,
    ks_datagen.py:5-89, started from u0 (physical space).  Returns n
    snapshots taken every `stride` steps after `n_discard` steps (the first
    snapshot is the state after n_discard steps)."""
    M2 = M // 2 + 1
    Mext = 3 * M // 2
    freq = 2 * np.pi / L * np.arange(M2)
    gg = -0.5j * Mext / M * freq
    ldiag = freq ** 2 - freq ** 4
    e1 = np.exp(dt * ldiag)
    e2 = np.exp(dt / 2 * ldiag)
    J = 16
    rvec = np.exp(1j * np.pi * np.linspace(0.5 / J, (J - 0.5) / J, num=J))
    Z = dt * ldiag.reshape(M2, 1) + rvec.reshape(1, J)
    f0 = dt * np.real(np.mean((np.exp(0.5 * Z) - 1) / Z, axis=-1))
    f1 = dt * np.real(np.mean((-4 - Z + np.exp(Z) * (4 - 3 * Z + Z ** 2)) / Z ** 3, axis=-1))
    f2 = dt * np.real(np.mean(2 * (2 + Z + np.exp(Z) * (-2 + Z)) / Z ** 3, axis=-1))
    f3 = dt * np.real(np.mean((-4 - 3 * Z - Z ** 2 + np.exp(Z) * (4 - Z)) / Z ** 3, axis=-1))

    def nl(v):
        return gg * np.fft.rfft(np.fft.irfft(v, n=Mext) ** 2)[:M2]

    v = np.fft.rfft(u0)

    def step(v):
        Nu = nl(v)
        a = e2 * v + f0 * Nu
        Na = nl(a)
        b = e2 * v + f0 * Na
        Nb = nl(b)
        c = e2 * a + f0 * (2 * Nb - Nu)
        Nc = nl(c)
        return e1 * v + f1 * Nu + f2 * (Na + Nb) + f3 * Nc

    for _ in range(n_discard):
        v = step(v)
    out = np.empty((n, M))
    out[0] = np.fft.irfft(v, n=M)
    for i in range(1, n):
        for _ in range(stride):
            v = step(v)
        out[i] = np.fft.irfft(v, n=M)
    return out


def gauss_ic(M, seed, scale=0.1):
    """0.1 N(0,1) per grid point: Box-Muller on Philox (GAUSS stream)."""
    u = np.empty(M)
    for p in range(0, M, 2):
        o = philox([p // 2, 0, STREAM_GAUSS, 0], [seed & 0xffffffff, seed >> 32])
        a = ((o[1] << 32) | o[0]) >> 11
        b = ((o[3] << 32) | o[2]) >> 11
        u1 = (a + 1) * 2.0 ** -53
        u2 = b * 2.0 ** -53
        r = np.sqrt(-2.0 * np.log(u1))
        u[p] = scale * r * np.cos(2 * np.pi * u2)
        if p + 1 < M:
            u[p + 1] = scale * r * np.sin(2 * np.pi * u2)
    return u
