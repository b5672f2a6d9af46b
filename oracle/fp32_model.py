"""Rounding model of the mixed-precision ARPC factor (test infrastructure).

Only tests/ and bench.py read this module; the library never imports it.
It restates the F-form blocked Cholesky with a fixed pivot sequence
(ks_arpclm.cu:929-995 with the pivots of a finished run) in NumPy and lets
each FP32 step of the library's mixed path be switched on alone, so that
the factor's distance from the all-FP64 factor can be attributed to the
step that causes it:

  kernel   K(:,S_i) as the mixed downdate forms it: FP32 points, FP32
           distance dot products, FP64 norms, exp rounded to FP32
  input    K(:,S_i) = fl32(exact K): every K value rounded once to FP32,
           all arithmetic FP64 (the floor of any method whose kernel values
           carry FP32 rounding)
  store    F rounded to FP32 after every round, arithmetic FP64
  gemm     the overlap F F(S_i,:)^T and T L_c^-T as FP32 products of FP32
           operands (float32 matmul), F stored FP32
  mixed    all of the above together (the library's mixed path)
  promote  as mixed, but the overlap accumulated in FP32 over chunks of
           `chunk` columns and the chunks summed in FP64
  late64   mixed until the residual trace falls below `switch`, FP64 from
           then on (F kept in FP64 from the switch; earlier columns as the
           mixed rounds left them)

The pivot block H = K(S',S') - F(S',:) F(S',:)^T and its Cholesky factor
L_c stay FP64 in every variant, as in the library (they are formed from the
stored F).  The pivots are injected as the library injects them
(ks_arpclm.cu:300-408 with every uniform test passed): each round proposes
the next `block` entries of the list and accepts, in order, those whose
Schur-complement diagonal is still above tol.
"""
import numpy as np

VARIANTS = ("fp64", "kernel", "input", "store", "gemm", "mixed", "promote", "late64")


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _kernel_cols(X64, X32, nrm32, S, eps, mode):
    """K(:, S) in the given rounding mode (n x |S|, FP64 array)."""
    if mode == "kernel":
        dot = (X32 @ X32[S].T).astype(np.float64)      # FP32 accumulation
        D = nrm32[:, None] + nrm32[S][None, :] - 2.0 * dot
        D = np.maximum(D, 0.0)
        K = _f32(np.exp(-D / eps))
        K[S, np.arange(len(S))] = 1.0                    # a pivot's own row is exactly 1
        return K
    sq = np.sum(X64 * X64, axis=1)
    D = np.maximum(sq[:, None] + sq[S][None, :] - 2.0 * (X64 @ X64[S].T), 0.0)
    K = np.exp(-D / eps)
    K[S, np.arange(len(S))] = 1.0
    return _f32(K) if mode == "input" else K


def _inject_block(H, tol):
    """Accepted positions and L_c (lower, FP64) of the injected proposals:
    proposal i is accepted when its current Schur diagonal exceeds tol."""
    H = np.array(H, dtype=np.float64)
    b = H.shape[0]
    acc = []
    cols = []
    for i in range(b):
        hii = H[i, i]
        if not hii > tol:
            continue
        l = H[:, i] / np.sqrt(hii)
        l[:i] = 0.0
        H -= np.outer(l, l)
        acc.append(i)
        cols.append(l)
    if not acc:
        return [], np.zeros((0, 0))
    L = np.array(cols).T[acc]          # rows: accepted proposals, columns: steps
    return acc, L


def factor(X, eps, pivots, rank, block=128, variant="fp64", chunk=256, switch=1e-3, tol=1e-10):
    """Blocked Cholesky factor F (n x rank) of exp(-|x-y|^2/eps) on the
    pivot list `pivots` injected `block` at a time, with the rounding of
    `variant`.  Returns (F, accepted pivots in order, trace errors per
    round)."""
    if variant not in VARIANTS:
        raise ValueError(variant)
    X64 = np.asarray(X, dtype=np.float64)
    X64 = X64 - X64.mean(axis=0)                 # the library centres the points
    X32f = X64.astype(np.float32)
    X32 = X32f.astype(np.float64)
    nrm32 = np.sum(X32 * X32, axis=1)
    n = X64.shape[0]
    F = np.zeros((n, rank))
    d = np.ones(n)
    errs, taken = [], []
    Nr, pos = 0, 0
    pivots = np.asarray(pivots, dtype=np.int64)
    while Nr < rank and pos < len(pivots) and d.max() > tol:
        Sp = pivots[pos:pos + block]
        pos += len(Sp)
        active = variant
        if variant == "late64" and d.sum() / n < switch:
            active = "fp64"
        kmode = {"kernel": "kernel", "input": "input", "mixed": "kernel", "promote": "kernel"}.get(active)
        if variant == "late64" and active != "fp64":
            kmode = "kernel"
        # pivot block in FP64 from the stored factor and FP64 kernel values
        Fo = F[:, :Nr]
        Kpp = _kernel_cols(X64[Sp], X64[Sp], None, np.arange(len(Sp)), eps, "exact")
        H = Kpp - Fo[Sp] @ Fo[Sp].T
        acc, Lc = _inject_block(H, tol)
        k = min(len(acc), rank - Nr)
        if k == 0:
            errs.append(d.sum() / n)
            continue
        acc, Lc = acc[:k], Lc[:k, :k]
        S = Sp[acc]
        Li = np.linalg.inv(Lc)
        Kc = _kernel_cols(X64, X32f if kmode == "kernel" else X64, nrm32, S, eps, kmode or "exact")
        FS = Fo[S]
        fp32_gemm = active in ("gemm", "mixed", "promote") or (variant == "late64" and active != "fp64")
        if fp32_gemm:
            F32 = Fo.astype(np.float32)
            FS32 = FS.astype(np.float32)
            if active == "promote":
                ov = np.zeros((n, k))
                for c0 in range(0, Nr, chunk):
                    ov += (F32[:, c0:c0 + chunk] @ FS32[:, c0:c0 + chunk].T).astype(np.float64)
                T = _f32(Kc - ov)
            else:
                ov = (F32 @ FS32.T) if Nr > 0 else np.zeros((n, k), np.float32)
                T = (Kc.astype(np.float32) - ov).astype(np.float32)
            G = (np.asarray(T, np.float32) @ Li.T.astype(np.float32)).astype(np.float64)
        else:
            T = Kc - Fo @ FS.T
            G = T @ Li.T
        store32 = active in ("store", "gemm", "mixed", "promote") or (variant == "late64" and active != "fp64")
        if store32:
            G = _f32(G)
        F[:, Nr:Nr + k] = G
        d = np.maximum(d - np.sum(G * G, axis=1), 0.0)
        Nr += k
        taken.extend(S.tolist())
        errs.append(d.sum() / n)
    return F[:, :Nr], np.array(taken, dtype=np.int64), np.array(errs)


def rel_frobenius(A, B):
    return float(np.linalg.norm(A - B) / np.linalg.norm(B))
