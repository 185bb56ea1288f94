"""The machine-peak measurements bench.py reports beside the datasheet
figures (prandtl_peak_fma, prandtl_peak_copy): plausible magnitudes on an
MI355X (datasheet: 78.6 TFLOP/s FP64 vector, 157.3 FP32 vector counting
packed FP32, 8 TB/s HBM), and unpacked FP32 FMA at the FP64 rate, which the
mixed form's bound relies on."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_measured_peaks(cuda):
    import bench
    p = bench.measured_peaks(cuda)
    print("measured peaks:", {k: round(v, 1) for k, v in p.items()}, flush=True)
    assert 20.0 < p["fp64_tflops"] <= 80.0
    assert 20.0 < p["fp32_tflops"] <= 160.0
    # no packed FP32 in this build: one FMA per lane per instruction either way
    assert 0.75 < p["fp32_tflops"] / p["fp64_tflops"] < 1.33
    assert 1500.0 < p["copy_GBps_1GiB"] <= 8000.0
    assert p["copy_GBps_27MB"] > 1000.0


def test_peak_entry_points_refuse_bad_arguments(cuda):
    from prandtl_amd._lib import lib, PRANDTL_ERR_ARG
    L = lib()
    t = ctypes.c_double()
    assert L.prandtl_peak_fma(16, ctypes.c_void_p(1), ctypes.byref(t), None) == PRANDTL_ERR_ARG
    assert L.prandtl_peak_copy(ctypes.c_void_p(1), ctypes.c_void_p(2), 24, 1, ctypes.byref(t), None) == PRANDTL_ERR_ARG
    assert L.prandtl_peak_copy(ctypes.c_void_p(1), ctypes.c_void_p(2), 32, 0, ctypes.byref(t), None) == PRANDTL_ERR_ARG
