"""nsn_measure_peaks: the measured ceilings are positive and below the
datasheet figures (78.6 / 157.3 TFLOP/s dense FP64 / FP32 MFMA, 8 TB/s HBM),
with 10 % slack for clock variation."""
import ctypes as C

import pytest

from nelson_amd import capi

pytestmark = pytest.mark.gpu


def test_measured_peaks_plausible():
    f64, f32, bw = C.c_double(0), C.c_double(0), C.c_double(0)
    capi.check(capi.lib().nsn_measure_peaks(0, C.byref(f64), C.byref(f32), C.byref(bw)))
    print(f"\nFP64 MFMA {f64.value:.1f} TFLOP/s, FP32 MFMA {f32.value:.1f} TFLOP/s, copy {bw.value:.0f} GB/s")
    assert 1.0 < f64.value < 1.1 * 78.6
    assert 1.0 < f32.value < 1.1 * 157.3
    assert 100.0 < bw.value < 1.1 * 8000.0


def test_measured_peaks_bad_device():
    assert capi.lib().nsn_measure_peaks(99, None, None, None) == capi.NSN_EINVAL
