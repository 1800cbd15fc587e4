"""bench.py's contract pieces that run without a GPU: the one-GPU guard, the
kernel names and HBM traffic figures it attaches to the roofline (the committed
rocprofv3 summary holds one figure per pole-kernel form, and the line must carry
the figure of the form the autotuner kept), and the roofline arithmetic."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("tag", ["fp32", "fp32_fused", "fp64"])
def test_traffic_figures_exist_for_every_form(tag):
    t = bench.measured_traffic(tag)
    assert t is not None and 5e7 < t < 2e9, (tag, t)


def test_fused_form_moves_less_than_the_slot_form():
    # the fused form keeps the pole sums in LDS; the slot form writes a float pair per pole
    assert bench.measured_traffic("fp32_fused") < 0.6 * bench.measured_traffic("fp32")


def test_kernel_names_follow_the_kept_shape():
    # bucket tuples: (nmax, nwin, T64, G64, wide64, T32, G32, wide32); wide 2 = matrix cores
    fused = (84, 512, 6, 4, 2, 6, 0, 2)
    slot = (84, 512, 6, 4, 2, 6, 1, 2)
    assert "window-fused" in bench.kernel_name(fused, True)
    assert bench.kernel_name(slot, True) == "k_pole_mfma<float, tiles 6, waves 1>"
    assert bench.kernel_name(fused, False) == "k_pole_mfma<double, tiles 6, waves 4>"
    assert bench.kernel_name((228, 512, 0, 0, 1, 15, 8, 2), False).startswith("k_pole_big<double")


def test_roofline_fraction_is_flops_over_time_over_peak():
    r = bench.roofline((4, 2.0), 51.7e9, bench.PEAK_FP32, "k", 2.0e8)
    assert r["launch_ms"] == 0.5
    assert abs(r["achieved"] - 51.7e9 / 0.5e-3 / 1e12) < 1e-9
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-12
    assert r["peak"] == 157.3 and r["traffic"] == 2.0e8
    assert bench.roofline((0, 0.0), 51.7e9, bench.PEAK_FP32, "k", None) is None
