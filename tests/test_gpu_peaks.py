"""GPU test of kevd_measure_peaks: every measured rate is positive and lies
between a floor well under MI355X's sustained rates and its published peak
(plus 10% for clock variation), so a probe that lost its work to the
compiler (a rate far above the peak) or ran serialised (far below) fails;
out-of-range copy sizes are rejected."""
import pytest

pytestmark = pytest.mark.gpu


def test_measure_peaks_plausible(gpu_ctx):
    pk = gpu_ctx.measure_peaks(copy_bytes=1 << 30)
    print("peaks", pk, flush=True)
    bounds = {"mfma_f32_tflops": (40.0, 157.3), "mfma_f32_32_tflops": (40.0, 157.3),
              "mfma_f64_tflops": (20.0, 78.6), "exp_f32_per_s": (1e12, 80e12),
              "hbm_copy_gbs": (1500.0, 8000.0), "hbm_read_gbs": (1500.0, 8000.0)}
    for k, (lo, hi) in bounds.items():
        assert lo < pk[k] < 1.1 * hi, (k, pk[k])


def test_measure_peaks_rejects_bad_size(gpu_ctx):
    from kernel_evd_amd import KevdError
    with pytest.raises(KevdError):
        gpu_ctx.measure_peaks(copy_bytes=1 << 10)
