"""ct_measure_peaks: the device's sustained matrix-core, FP64 vector and HBM
rates, against which bench.py also states each kernel's roofline fraction.

The probes must land below the published MI355X peaks (bf16 dense MFMA
2.5 PFLOP/s, FP64 MFMA and vector 78.6 TFLOP/s, HBM 8 TB/s) and well above
what a probe with a broken flop count or a serialised loop would show (a
quarter of the peak or less).
"""
import pytest

pytestmark = pytest.mark.gpu


def test_peaks_plausible():
    from chemtoolscuda_amd import capi

    p = capi.measure_peaks()
    print(p)
    assert 0.6 * 2500 < p["bf16_mfma_tflops"] <= 1.02 * 2500
    assert 0.6 * 78.6 < p["f64_mfma_tflops"] <= 1.02 * 78.6
    assert 0.5 * 78.6 < p["f64_valu_tflops"] <= 1.02 * 78.6
    assert 0.4 * 8000 < p["hbm_copy_gbs"] <= 8000
