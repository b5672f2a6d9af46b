"""ctypes binding of libkevd.so (the C ABI declared in include/kevd.h).

The library is loaded from this package directory only; there is no fallback
implementation.  A missing or unloadable library raises at first use.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkevd.so")

KEVD_OK, KEVD_EINVAL, KEVD_ENONPOS_DEGREE, KEVD_EPOTRF, KEVD_ENONPOS_Q, KEVD_EEIG, \
    KEVD_EHIP, KEVD_ENOMEM = range(8)
KEVD_FP64, KEVD_MIXED = 0, 1
KEVD_NORM_REF, KEVD_NORM_SINKHORN = 0, 1

i32, i64, u64, dbl = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER


class ArpcParams(C.Structure):
    _fields_ = [("eps", dbl), ("rank", i64), ("block", i32), ("seed", u64), ("tol", dbl),
                ("prec", C.c_int), ("max_rounds", i32), ("inject_pivots", P(i64)),
                ("n_inject", i64), ("regenerate", i32)]


class ArpcResult(C.Structure):
    _fields_ = [("rank", i64), ("rounds", i32), ("pivots", P(i64)), ("accepted", P(i32)),
                ("trace_err", P(dbl)), ("rounds_cap", i32)]


class EvdParams(C.Structure):
    _fields_ = [("k_out", i32), ("norm", i32), ("sinkhorn_tol", dbl),
                ("sinkhorn_max_iter", i32), ("prec", C.c_int), ("all_eigenvalues", i32)]


class EvdResult(C.Structure):
    _fields_ = [("lam", P(dbl)), ("U", P(C.c_float)), ("rowsum_dev", dbl), ("shift", dbl),
                ("min_degree", dbl), ("sinkhorn_iters", i32)]


class KsParams(C.Structure):
    _fields_ = [("L", dbl), ("M", i32), ("dt", dbl), ("n_discard", i64), ("stride", i32),
                ("seed", u64), ("ic_scale", dbl)]


class Timings(C.Structure):
    _fields_ = [("arpc_ms", dbl), ("evd_ms", dbl), ("downdate_ms", dbl),
                ("downdate_launches", i64), ("downdate_flop", dbl), ("gram_ms", dbl),
                ("gram_launches", i64), ("gram_flop", dbl), ("eig_ms", dbl),
                ("sample_ms", dbl), ("rejection_ms", dbl), ("matvec_ms", dbl),
                ("dense_ms", dbl), ("u_ms", dbl), ("eig_iters", i32), ("eig_householder", i32),
                ("eig_full_fallback", i32), ("eig_batched", i32),
                ("eig_orth_passes", i32), ("sinkhorn_batched", i32)]


class Peaks(C.Structure):
    _fields_ = [("mfma_f32_tflops", dbl), ("mfma_f32_32_tflops", dbl), ("mfma_f64_tflops", dbl),
                ("exp_f32_per_s", dbl), ("hbm_copy_gbs", dbl), ("hbm_read_gbs", dbl),
                ("mfma_bf16_tflops", dbl), ("mfma_bf16_32_tflops", dbl)]


# name -> (restype, argtypes); every function declared in include/kevd.h
SIGNATURES = {
    "kevd_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
    "kevd_destroy": (None, [C.c_void_p]),
    "kevd_last_error": (C.c_char_p, [C.c_void_p]),
    "kevd_status_string": (C.c_char_p, [C.c_int]),
    "kevd_set_points": (C.c_int, [C.c_void_p, P(dbl), i64, i32]),
    "kevd_set_points_f32": (C.c_int, [C.c_void_p, P(C.c_float), i64, i32]),
    "kevd_median_sqdist": (C.c_int, [C.c_void_p, u64, i64, P(dbl)]),
    "kevd_kernel_block": (C.c_int, [C.c_void_p, P(i64), i64, P(i64), i64, dbl, C.c_int, P(dbl)]),
    "kevd_arpc": (C.c_int, [C.c_void_p, P(ArpcParams), P(ArpcResult)]),
    "kevd_get_factor": (C.c_int, [C.c_void_p, C.c_void_p, i64, C.c_int]),
    "kevd_set_factor": (C.c_int, [C.c_void_p, C.c_void_p, i64, i64, i64, C.c_int]),
    "kevd_get_linv": (C.c_int, [C.c_void_p, P(dbl)]),
    "kevd_evd": (C.c_int, [C.c_void_p, P(EvdParams), P(EvdResult)]),
    "kevd_get_u": (C.c_int, [C.c_void_p, P(C.c_float)]),
    "kevd_set_dump": (C.c_int, [C.c_void_p, i32]),
    "kevd_get_intermediate": (C.c_int, [C.c_void_p, C.c_char_p, P(dbl), P(i64)]),
    "kevd_bandwidth_sweep": (C.c_int, [C.c_void_p, P(i64), i64, P(dbl), i32, C.c_int, P(dbl)]),
    "kevd_bandwidth_sweep_noalloc": (C.c_int, [C.c_void_p, P(i64), i64, P(dbl), i32, C.c_int, P(dbl)]),
    "kevd_bandwidth_post": (C.c_int, [P(dbl), P(dbl), i32, P(dbl), P(dbl), P(dbl), P(dbl)]),
    "kevd_ks_generate": (C.c_int, [P(KsParams), i64, P(dbl)]),
    "kevd_delay_embed": (C.c_int, [P(dbl), i64, i32, i32, P(dbl)]),
    "kevd_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "kevd_get_timings": (C.c_int, [C.c_void_p, P(Timings)]),
    "kevd_ensemble_arpc": (C.c_int, [P(C.c_void_p), i32, P(ArpcParams), P(ArpcResult), P(i32)]),
    "kevd_ensemble_evd": (C.c_int, [P(C.c_void_p), i32, P(EvdParams), P(EvdResult), i32]),
    "kevd_autotune_block": (C.c_int, [C.c_void_p, P(ArpcParams), P(i32), i32, P(i32), P(dbl)]),
    "kevd_autotune_tiles": (C.c_int, [C.c_void_p, P(i32), P(dbl)]),
    "kevd_set_gram_tile": (C.c_int, [C.c_void_p, i32]),
    "kevd_measure_peaks": (C.c_int, [C.c_void_p, i64, P(Peaks)]),
}

_lib = None


def lib():
    """Load libkevd.so from the package directory (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkevd.so is not built; run __graft_entry__.build() "
                               "or python -m kernel_evd_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class KevdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("kevd status %d: %s" % (status, msg))
        self.status = status
