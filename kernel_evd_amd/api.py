"""Python face of libkevd: a thin object wrapper over the C ABI.

It mirrors the three reference programs as calls on one GPU context:
``arpc`` (ks_arpclm.cu main), ``evd`` (ks_kevd.cu main) and
``bandwidth_sweep`` + ``bandwidth_post`` (ks_bandwidth.cu main).  All
arithmetic runs in libkevd on the GPU; this module only moves arrays.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import KevdError, KEVD_FP64, KEVD_MIXED, KEVD_NORM_REF, KEVD_NORM_SINKHORN

PREC = {"fp64": KEVD_FP64, "mixed": KEVD_MIXED, KEVD_FP64: KEVD_FP64, KEVD_MIXED: KEVD_MIXED}


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class ArpcOutput:
    def __init__(self, rank, rounds, pivots, accepted, trace_err):
        self.rank = rank
        self.rounds = rounds
        self.pivots = pivots
        self.accepted = accepted
        self.trace_err = trace_err


class EvdOutput:
    def __init__(self, lam, U, rowsum_dev, shift, min_degree, sinkhorn_iters):
        self.lam = lam
        self.U = U
        self.rowsum_dev = rowsum_dev
        self.shift = shift
        self.min_degree = min_degree
        self.sinkhorn_iters = sinkhorn_iters


class Context:
    """One GPU context (device memory, stream).  Use as a context manager."""

    def __init__(self, device=0):
        self._lib = L.lib()
        h = C.c_void_p()
        s = self._lib.kevd_create(device, C.byref(h))
        if s != 0:
            raise KevdError(s, self._lib.kevd_status_string(s).decode())
        self._h = h
        self.n = 0
        self.dim = 0
        self._k_out = 0
        self._rank = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kevd_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, s):
        if s != 0:
            raise KevdError(s, self._lib.kevd_last_error(self._h).decode())

    # ------------------------------------------------------------ points
    def set_points(self, X):
        X = np.asarray(X)
        if X.ndim != 2:
            raise ValueError("X must be n x dim")
        if X.dtype == np.float32:
            X = np.ascontiguousarray(X)
            self._check(self._lib.kevd_set_points_f32(self._h, _p(X, C.c_float), X.shape[0], X.shape[1]))
        else:
            X = np.ascontiguousarray(X, dtype=np.float64)
            self._check(self._lib.kevd_set_points(self._h, _p(X, C.c_double), X.shape[0], X.shape[1]))
        self.n, self.dim = X.shape
        self._rank = 0

    def median_sqdist(self, seed=0, nsample=0):
        out = C.c_double()
        self._check(self._lib.kevd_median_sqdist(self._h, seed, nsample, C.byref(out)))
        return out.value

    def kernel_block(self, rows, cols, eps, prec="fp64"):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        out = np.empty((rows.size, cols.size), dtype=np.float64)
        self._check(self._lib.kevd_kernel_block(self._h, _p(rows, C.c_int64), rows.size,
                                                _p(cols, C.c_int64), cols.size, eps,
                                                PREC[prec], _p(out, C.c_double)))
        return out

    # -------------------------------------------------------------- ARPC
    def arpc(self, eps, rank, block=128, seed=0, tol=1e-10, prec="mixed", max_rounds=0,
             inject_pivots=None, regenerate=False):
        """kevd_arpc.  regenerate=True runs the LM form (no factor stored;
        L^-1 through get_linv)."""
        p = L.ArpcParams()
        p.regenerate = 1 if regenerate else 0
        p.eps = eps
        p.rank = rank
        p.block = block
        p.seed = seed
        p.tol = tol
        p.prec = PREC[prec]
        mr = max_rounds if max_rounds > 0 else 4 * ((rank + block - 1) // block) + 64
        p.max_rounds = mr
        inj = None
        if inject_pivots is not None:
            inj = np.ascontiguousarray(inject_pivots, dtype=np.int64)
            p.inject_pivots = _p(inj, C.c_int64)
            p.n_inject = inj.size
        piv = np.zeros(rank, dtype=np.int64)
        acc = np.zeros(mr, dtype=np.int32)
        err = np.zeros(mr, dtype=np.float64)
        r = L.ArpcResult()
        r.pivots = _p(piv, C.c_int64)
        r.accepted = _p(acc, C.c_int32)
        r.trace_err = _p(err, C.c_double)
        r.rounds_cap = mr
        self._check(self._lib.kevd_arpc(self._h, C.byref(p), C.byref(r)))
        self._rank = 0 if regenerate else r.rank
        nr = r.rounds
        return ArpcOutput(r.rank, nr, piv[:r.rank].copy(), acc[:nr].copy(), err[:nr].copy())

    def get_factor(self, rank=None, as_prec="fp64"):
        """The factor held by the context (n x rank); rank defaults to the
        realised rank and may not exceed it."""
        rank = self._rank if rank is None else rank
        if rank != self._rank:
            raise KevdError(1, "get_factor: the context holds a factor of rank %d, not %d" % (self._rank, rank))
        dt = np.float32 if PREC[as_prec] == KEVD_MIXED else np.float64
        F = np.empty((self.n, rank), dtype=dt)
        self._check(self._lib.kevd_get_factor(self._h, F.ctypes.data_as(C.c_void_p), rank, PREC[as_prec]))
        return F

    def set_factor(self, F):
        F = np.asarray(F)
        prec = KEVD_MIXED if F.dtype == np.float32 else KEVD_FP64
        F = np.ascontiguousarray(F, dtype=np.float32 if prec == KEVD_MIXED else np.float64)
        self._check(self._lib.kevd_set_factor(self._h, F.ctypes.data_as(C.c_void_p), F.shape[0],
                                              F.shape[1], F.shape[1], prec))
        self._rank = F.shape[1]

    def get_linv(self, rank):
        out = np.empty(rank * rank, dtype=np.float64)
        self._check(self._lib.kevd_get_linv(self._h, _p(out, C.c_double)))
        return out.reshape(rank, rank, order="F")

    # -------------------------------------------------------------- kEVD
    def evd(self, rank, k_out=100, norm="ref", prec="mixed", sinkhorn_tol=1e-12,
            sinkhorn_max_iter=1000, want_U=True, all_eigenvalues=False):
        p = L.EvdParams()
        p.k_out = k_out
        p.norm = KEVD_NORM_SINKHORN if norm == "sinkhorn" else KEVD_NORM_REF
        p.sinkhorn_tol = sinkhorn_tol
        p.sinkhorn_max_iter = sinkhorn_max_iter
        p.prec = PREC[prec]
        p.all_eigenvalues = 1 if all_eigenvalues else 0
        lam = np.zeros(rank, dtype=np.float64)
        r = L.EvdResult()
        r.lam = _p(lam, C.c_double)
        U = None
        if want_U:
            U = np.zeros((self.n, k_out), dtype=np.float32)
            r.U = _p(U, C.c_float)
        self._check(self._lib.kevd_evd(self._h, C.byref(p), C.byref(r)))
        self._k_out = k_out
        return EvdOutput(lam, U, r.rowsum_dev, r.shift, r.min_degree, r.sinkhorn_iters)

    def get_u(self):
        """Eigenvectors of the last evd call (n x k_out FP32, columns ascending)."""
        k = self._k_out
        U = np.empty((self.n, k), dtype=np.float32)
        self._check(self._lib.kevd_get_u(self._h, _p(U, C.c_float)))
        return U

    # The reference's dump datasets (ks_kevd.cu:531-542 and its later
    # h5add_double sites), in their order there.
    DUMP_NAMES = ("c1", "p1", "dtil", "cq", "pq", "Atil", "A", "R", "G", "qtil", "Btil", "W", "V", "lam", "Z")

    def set_dump(self, on=True):
        """Record the kEVD's intermediates on the next evd calls
        (kevd_set_dump)."""
        self._check(self._lib.kevd_set_dump(self._h, 1 if on else 0))

    def intermediate(self, name):
        """One recorded intermediate, FP64; matrices come back r x r (V, Z:
        r x columns) from the column-major record."""
        cnt = C.c_int64(0)
        self._check(self._lib.kevd_get_intermediate(self._h, name.encode(), None, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float64)
        self._check(self._lib.kevd_get_intermediate(self._h, name.encode(), _p(out, C.c_double), C.byref(cnt)))
        if name in ("Atil", "A", "R", "G", "Btil", "W", "V", "Z"):
            r = C.c_int64(0)  # the rank, from a length-r entry
            self._check(self._lib.kevd_get_intermediate(self._h, b"c1", None, C.byref(r)))
            return out.reshape(r.value, out.size // r.value, order="F")
        return out

    def intermediates(self):
        """Every recorded intermediate by name."""
        return {name: self.intermediate(name) for name in self.DUMP_NAMES}

    # --------------------------------------------------------- bandwidth
    def bandwidth_sweep(self, rows, eps, prec="mixed", noalloc=False):
        """kevd_bandwidth_sweep; noalloc=True calls the allocation-free
        variant (kevd_bandwidth_sweep_noalloc)."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        S = np.zeros(eps.size, dtype=np.float64)
        f = self._lib.kevd_bandwidth_sweep_noalloc if noalloc else self._lib.kevd_bandwidth_sweep
        self._check(f(self._h, _p(rows, C.c_int64), rows.size, _p(eps, C.c_double), eps.size, PREC[prec],
                      _p(S, C.c_double)))
        return S

    def autotune_block(self, eps, rank, candidates=(32, 64, 128), seed=0, tol=1e-10, prec="mixed",
                       measure=False):
        """Block size with the best factor columns/s for this point set
        (kevd_autotune_block).  measure=True times every candidate and
        returns the rates too; otherwise a cached choice may be returned."""
        p = L.ArpcParams()
        p.eps = eps
        p.rank = rank
        p.block = 1
        p.seed = seed
        p.tol = tol
        p.prec = PREC[prec]
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        best = C.c_int32()
        rates = np.zeros(cand.size) if measure else None
        self._check(self._lib.kevd_autotune_block(self._h, C.byref(p), _p(cand, C.c_int32), cand.size,
                                                  C.byref(best), _p(rates, C.c_double) if measure else None))
        self._rank = rank  # the context holds the last timed run's factor
        return (best.value, rates) if measure else best.value

    def autotune_tiles(self, measure=False):
        """The bf16x6 Gram's output tile for this context's factor
        (kevd_autotune_tiles): 256, 128, or 0 when there is no choice.
        measure=True times both (and returns their ms as well)."""
        t = C.c_int32(0)
        ms = np.zeros(2) if measure else None
        self._check(self._lib.kevd_autotune_tiles(self._h, C.byref(t), _p(ms, C.c_double) if measure else None))
        return (t.value, ms) if measure else t.value

    def set_gram_tile(self, tile):
        """Use the given bf16x6 Gram tile (0, 128 or 256) without measuring."""
        self._check(self._lib.kevd_set_gram_tile(self._h, int(tile)))

    def measure_peaks(self, copy_bytes=4 << 30):
        """Sustained MFMA (FP32 16x16 / 32x32, FP64), v_exp_f32 and HBM copy
        rates of this context's device (kevd_measure_peaks)."""
        pk = L.Peaks()
        self._check(self._lib.kevd_measure_peaks(self._h, int(copy_bytes), C.byref(pk)))
        return {name: getattr(pk, name) for name, _ in L.Peaks._fields_}

    # ------------------------------------------------------------ timing
    def set_profiling(self, on=True):
        self._check(self._lib.kevd_set_profiling(self._h, 1 if on else 0))

    def timings(self):
        t = L.Timings()
        self._check(self._lib.kevd_get_timings(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in L.Timings._fields_}


def ensemble_arpc(contexts, eps, ranks, blocks, seeds=None, tol=1e-10, prec="mixed"):
    """Batched ARPC over several contexts (kevd_ensemble_arpc).  eps, ranks,
    blocks (and seeds) are per-problem sequences.  Returns (list of
    ArpcOutput, used_graph)."""
    lib = L.lib()
    P = len(contexts)
    seeds = [0] * P if seeds is None else list(seeds)
    ps = (L.ArpcParams * P)()
    rs = (L.ArpcResult * P)()
    keep = []
    for i, ctx in enumerate(contexts):
        ps[i].eps = eps[i]
        ps[i].rank = ranks[i]
        ps[i].block = blocks[i]
        ps[i].seed = seeds[i]
        ps[i].tol = tol
        ps[i].prec = PREC[prec]
        mr = 4 * ((ranks[i] + blocks[i] - 1) // blocks[i]) + 64
        ps[i].max_rounds = mr
        piv = np.zeros(ranks[i], dtype=np.int64)
        acc = np.zeros(mr, dtype=np.int32)
        err = np.zeros(mr, dtype=np.float64)
        keep.append((piv, acc, err))
        rs[i].pivots = _p(piv, C.c_int64)
        rs[i].accepted = _p(acc, C.c_int32)
        rs[i].trace_err = _p(err, C.c_double)
        rs[i].rounds_cap = mr
    hs = (C.c_void_p * P)(*[c._h.value for c in contexts])
    used = C.c_int32()
    st = lib.kevd_ensemble_arpc(hs, P, ps, rs, C.byref(used))
    if st != 0:
        raise KevdError(st, lib.kevd_last_error(contexts[0]._h).decode())
    outs = []
    for i in range(P):
        contexts[i]._rank = rs[i].rank
        piv, acc, err = keep[i]
        nr = rs[i].rounds
        outs.append(ArpcOutput(rs[i].rank, nr, piv[:rs[i].rank].copy(), acc[:nr].copy(), err[:nr].copy()))
    return outs, bool(used.value)


def ensemble_evd(contexts, ranks, k_out=32, norm="sinkhorn", prec="mixed", sinkhorn_tol=1e-12,
                 sinkhorn_max_iter=1000, want_U=False, workers=8):
    """kevd_evd on every context concurrently (kevd_ensemble_evd)."""
    lib = L.lib()
    P = len(contexts)
    p = L.EvdParams()
    p.k_out = k_out
    p.norm = KEVD_NORM_SINKHORN if norm == "sinkhorn" else KEVD_NORM_REF
    p.sinkhorn_tol = sinkhorn_tol
    p.sinkhorn_max_iter = sinkhorn_max_iter
    p.prec = PREC[prec]
    p.all_eigenvalues = 0
    rs = (L.EvdResult * P)()
    lams, Us = [], []
    for i, ctx in enumerate(contexts):
        lam = np.zeros(ranks[i])
        lams.append(lam)
        rs[i].lam = _p(lam, C.c_double)
        U = None
        if want_U:
            U = np.zeros((ctx.n, k_out), dtype=np.float32)
            rs[i].U = _p(U, C.c_float)
        Us.append(U)
    hs = (C.c_void_p * P)(*[c._h.value for c in contexts])
    st = lib.kevd_ensemble_evd(hs, P, C.byref(p), rs, workers)
    if st != 0:
        raise KevdError(st, lib.kevd_last_error(contexts[0]._h).decode())
    for ctx in contexts:
        ctx._k_out = k_out
    return [EvdOutput(lams[i], Us[i], rs[i].rowsum_dev, rs[i].shift, rs[i].min_degree, rs[i].sinkhorn_iters)
            for i in range(P)]


def bandwidth_post(eps, S):
    """logS, slope, eopt, mdim of a sweep (ks_bandwidth.cu:609-642)."""
    lib = L.lib()
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    n = eps.size
    logS = np.zeros(n)
    slope = np.zeros(max(n - 2, 1))
    eopt = C.c_double()
    mdim = C.c_double()
    s = lib.kevd_bandwidth_post(_p(eps, C.c_double), _p(S, C.c_double), n, _p(logS, C.c_double),
                                _p(slope, C.c_double), C.byref(eopt), C.byref(mdim))
    if s != 0:
        raise KevdError(s, "bandwidth_post: need >= 3 positive values")
    return logS, slope[:n - 2], eopt.value, mdim.value


def ks_generate(n, L_dom=22.0, M=64, dt=0.25, n_discard=10000, stride=4, seed=0, ic_scale=0.1):
    """Kuramoto-Sivashinsky snapshots, n x M FP64 (ks_datagen.py:5-89)."""
    lib = L.lib()
    p = L.KsParams()
    p.L = L_dom
    p.M = M
    p.dt = dt
    p.n_discard = n_discard
    p.stride = stride
    p.seed = seed
    p.ic_scale = ic_scale
    out = np.empty((n, M), dtype=np.float64)
    s = lib.kevd_ks_generate(C.byref(p), n, _p(out, C.c_double))
    if s != 0:
        raise KevdError(s, "ks_generate: bad parameters")
    return out


def delay_embed(series, nq):
    """Delay embedding in the reference's point order: point g*N + t is the
    window series[t:t+nq, g] (ks_arpclm.cu:263-268)."""
    lib = L.lib()
    series = np.ascontiguousarray(series, dtype=np.float64)
    ndata, M = series.shape
    N = ndata - nq + 1
    out = np.empty((N * M, nq), dtype=np.float64)
    s = lib.kevd_delay_embed(_p(series, C.c_double), ndata, M, nq, _p(out, C.c_double))
    if s != 0:
        raise KevdError(s, "delay_embed: need 1 <= nq <= ndata")
    return out
