"""nspt_measure_peaks, the measured side of bench.py's roofline peaks: an
FP64 FMA-chain kernel and a 1 GiB device copy on the GPU.  The values must be
positive and no more than about the MI355X's published figures (78.6 TFLOP/s
FP64, 8 TB/s HBM), with room for clock variation."""
import pytest

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible():
    from prlgt2020_amd import nspt
    tflops, gbs = nspt.measure_peaks()
    print("FP64 FMA %.1f TFLOP/s, copy %.0f GB/s" % (tflops, gbs))
    assert 5.0 < tflops < 1.15 * 78.6
    assert 500.0 < gbs < 1.15 * 8000.0
