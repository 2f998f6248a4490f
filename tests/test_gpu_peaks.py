"""The measured-peak kernels (pir_measure_peak): each runs and reports a rate
within physical bounds of one MI355X (above a tenth of, and below twice, the
datasheet figure), and unknown peaks or sizes are refused."""
import ctypes

import pytest

import piranha_amd as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,lo,hi", [
    (P.PEAK_HBM_COPY, 800.0, 16000.0),     # GB/s, datasheet 8000
    (P.PEAK_HBM_READ, 800.0, 16000.0),
    (P.PEAK_HBM_WRITE, 800.0, 16000.0),
    (P.PEAK_MFMA_I8, 500.0, 10000.0),      # TOP/s, datasheet 5000 dense
    (P.PEAK_MFMA_I8_RANDOM, 500.0, 10000.0),
    (P.PEAK_VALU_U64, 100.0, 40000.0),     # G 64-bit multiply-adds / s
])
def test_peak_in_range(which, lo, hi):
    rate, ms = P.measure_peak(which, 1 << 28, reps=3)
    print(f"peak {which}: {rate:.1f} in {ms:.3f} ms", flush=True)
    assert lo < rate < hi and ms > 0


def test_peak_rejects_bad_arguments():
    L = P.lib()
    r, ms = ctypes.c_double(), ctypes.c_double()
    assert L.pir_measure_peak(99, 1 << 20, 1, ctypes.byref(r), ctypes.byref(ms), None) != 0
    assert L.pir_measure_peak(P.PEAK_HBM_COPY, 17, 1, ctypes.byref(r), ctypes.byref(ms), None) != 0
    assert L.pir_measure_peak(P.PEAK_HBM_COPY, 1 << 20, 0, ctypes.byref(r), ctypes.byref(ms), None) != 0
