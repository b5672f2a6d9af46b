"""The rounding model of the mixed factor (oracle/fp32_model.py), on the CPU.

With the pivot list of an FP64 run injected, the model switches on one FP32
step of the mixed path at a time and measures the factor's relative
Frobenius distance from the FP64 factor.  These tests pin the model (its
FP64 form is the oracle's factor; its FP32 kernel values are within FP32
rounding of the exact ones) and then the floors that decide whether a
factor tolerance of 1e-6 against FP64 is reachable (KS data, N = 8192,
rank 600, block 128, final trace error about 7e-4):

  input   1.4e-5   every K value rounded once to FP32, all else FP64
  store   5.5e-6   only F stored in FP32
  late64  8.1e-5   FP32 in the first round only, FP64 in every later one
  mixed   1.8e-4   the library's mixed path

The factor's late columns are L_c^-T times residuals that are small
differences of O(1) quantities, so an absolute error e in the residual of a
round whose pivots have residual diagonal ~h becomes e / sqrt(h) in F; FP32
rounding anywhere in the first rounds is carried into every later column.
"""
import numpy as np
import pytest

import fp32_model as fm
import oracle as orc

N, RANK, BLOCK = 8192, 600, 128


@pytest.fixture(scope="module")
def ks_case():
    from kernel_evd_amd import ks_generate
    X = ks_generate(N, L_dom=22.0, M=64, dt=0.25, n_discard=10000, stride=4, seed=0)
    eps = orc.median_sqdist(X, seed=0, nsample=1 << 20)
    ref = orc.arpc(X, eps, RANK, block=BLOCK, seed=0)
    assert ref.rank == RANK
    return X, eps, ref


def _err(case, variant, **kw):
    X, eps, ref = case
    F, piv, _ = fm.factor(X, eps, ref.pivots, RANK, block=BLOCK, variant=variant, **kw)
    assert np.array_equal(piv, ref.pivots), variant
    return fm.rel_frobenius(F, ref.F)


def test_model_fp64_is_the_oracle(ks_case):
    X, eps, ref = ks_case
    F, piv, errs = fm.factor(X, eps, ref.pivots, RANK, block=BLOCK, variant="fp64")
    np.testing.assert_array_equal(piv, ref.pivots)
    assert fm.rel_frobenius(F, ref.F) < 1e-10
    assert abs(errs[-1] - ref.trace_err[-1]) < 1e-10


def test_model_kernel_values_carry_fp32_rounding(ks_case):
    X, eps, ref = ks_case
    S = ref.pivots[:64]
    X64 = X - X.mean(axis=0)
    X32 = X64.astype(np.float32)
    nrm32 = np.sum(X32.astype(np.float64) ** 2, axis=1)
    K32 = fm._kernel_cols(X64, X32, nrm32, S, eps, "kernel")
    K = orc.kernel_block(X, np.arange(N), S, eps)
    diff = np.abs(K32 - K).max()
    assert 1e-9 < diff < 2e-6, diff  # the GPU kernel block bar of the mixed path


def test_floor_of_fp32_kernel_values(ks_case):
    # no method that consumes FP32-rounded kernel values gets closer
    assert _err(ks_case, "input") > 5e-6


def test_floor_of_fp32_factor_storage(ks_case):
    assert _err(ks_case, "store") > 2e-6


def test_fp64_after_the_first_round_does_not_reach_1e6(ks_case):
    # an adaptive switch to FP64 as early as after the first round
    assert _err(ks_case, "late64", switch=0.5) > 2e-5


def test_mixed_model_error(ks_case):
    e_mixed = _err(ks_case, "mixed")
    assert 5e-5 < e_mixed < 1e-3
    assert e_mixed > _err(ks_case, "input")
