"""CPU-side checks of the C ABI: libkevd.so loads, exports every function
include/kevd.h declares, the Python binding declares the same set, and the
host-only entry points behave (no GPU is touched here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as orc
from kernel_evd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "kevd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(kevd_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_symbols_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 18
    for n in sorted(names):
        assert hasattr(lib, n), "libkevd.so does not export %s" % n
    assert names == set(_lib.SIGNATURES), "binding and header disagree: %s" % (
        names ^ set(_lib.SIGNATURES))


def test_status_strings():
    L = _lib.lib()
    assert L.kevd_status_string(0) == b"ok"
    assert L.kevd_status_string(3).startswith(b"Cholesky")
    assert L.kevd_last_error(None) == b"null context"


def test_create_without_gpu_fails_cleanly():
    L = _lib.lib()
    h = C.c_void_p()
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    assert L.kevd_create(0, C.byref(h)) != 0
    assert not h.value


def test_ks_generator_matches_numpy_etdrk4():
    from kernel_evd_amd import ks_generate
    for M, L in [(64, 22.0), (32, 22.0), (128, 88.0)]:
        a = ks_generate(30, L_dom=L, M=M, n_discard=80, stride=4, seed=3)
        b = orc.ks_numpy(30, L=L, M=M, n_discard=80, stride=4, u0=orc.gauss_ic(M, 3))
        assert np.max(np.abs(a - b)) < 1e-10 * max(1.0, np.abs(b).max())


def test_ks_generator_checksum_fixture():
    """Golden checksum of the build's own generator output (the data set is
    defined by this generator: the trajectory is chaotic)."""
    from kernel_evd_amd import ks_generate
    X = ks_generate(2000, L_dom=22.0, M=64, n_discard=10000, stride=4, seed=0)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "ks_L22_M64_seed0_first2000_stats.npy"))
    got = np.array([X.mean(), X.std(), np.abs(X).max(), X[0, 0], X[-1, -1]])
    assert np.allclose(got, fx, rtol=1e-9, atol=1e-12)


def test_ks_generator_rejects_bad_params():
    from kernel_evd_amd import KevdError, ks_generate
    with pytest.raises(KevdError):
        ks_generate(10, M=30)
    with pytest.raises(KevdError):
        ks_generate(10, dt=-1.0)
