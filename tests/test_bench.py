"""CPU checks of bench.py's contract: more than one GPU is refused with a
message and exit status 2, and the roofline object is built from the live
profiler's families as documented (dominant kernel by time, achieved =
algorithmic bytes per launch / average launch time, the committed counter
traffic for that kernel, the copy-ceiling fraction, the FFT fractions)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_from_families():
    # {family: (ms, algorithmic bytes, launches)}
    prof = {
        "dp_stage": (100.0, 70.0e9, 10),        # 7 GB per launch in 10 ms: 700 GB/s
        "y_pass": (20.0, 8.0e9, 40),
        "z_r2c": (5.0, 2.0e9, 10),
        "y_tt_update_h": (30.0, 6.0e9, 12),
        "dp_finish": (0.0, 0.0, 0),             # no launches: left out
    }
    r = bench.roofline(prof, grid=256, copy_gbs=1000.0)
    assert r["kernel"] == "dp_stage"
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == bench.HBM_PEAK_GBS
    assert r["achieved"] == pytest.approx(700.0)
    assert r["frac"] == pytest.approx(700.0 / bench.HBM_PEAK_GBS, abs=1e-4)
    assert r["avg_launch_us"] == pytest.approx(10000.0)
    assert r["bytes_per_launch"] == pytest.approx(7.0e9)
    assert r["frac_of_copy_ceiling"] == pytest.approx(0.7)
    # the committed counter summary is for k_dp_stage at 256^3
    assert r["traffic"] == pytest.approx(7.0554e9, rel=1e-3)
    assert bench.roofline(prof, grid=128)["traffic"] is None
    # FFT fractions: all passes (y, z r2c, TT) and the plain ones (y, z r2c)
    fft_gbs = (8.0e9 + 2.0e9 + 6.0e9) / 0.055 / 1e9
    plain_gbs = (8.0e9 + 2.0e9) / 0.025 / 1e9
    assert r["fft_all_passes_GB_per_s"] == pytest.approx(fft_gbs, rel=1e-3)
    assert r["fft_plain_passes_GB_per_s"] == pytest.approx(plain_gbs, rel=1e-3)
    assert "dp_finish" not in r["kernels"]
    assert r["kernels"]["y_pass"]["avg_us"] == pytest.approx(500.0)
