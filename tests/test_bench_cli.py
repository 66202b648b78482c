"""CPU checks of bench.py's contract helpers (no GPU work)."""
import json
import os
import subprocess
import sys
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr
    assert out.stdout == ""


def test_roofline_counts_twenty_flops_per_interaction():
    r = bench.roofline(3.0e12, "f32", 1.5, traffic=123)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s"
    assert abs(r["achieved"] - 60.0) < 1e-9            # 3e12 * 20 / 1e12
    assert r["peak"] == 157.3 and abs(r["frac"] - 60.0 / 157.3) < 1e-12
    assert r["traffic"] == 123 and r["kernel_avg_ms"] == 1.5
    assert bench.roofline(1.0e12, "f64", 2.0)["peak"] == 78.6


def test_measured_traffic_reads_the_committed_profile():
    t = bench.measured_traffic()
    assert set(t) == {"f32r", "f32c", "f64"}
    # tens of MB per launch at N=65536 (partial sums of the split range)
    assert all(1e7 < v < 1e9 for v in t.values())


def test_default_f32_form_per_workload():
    assert bench.default_f32("single") == "f32r"
    assert bench.default_f32("large") == "f32r"
    assert bench.default_f32("batched") == "f32c"
    assert bench.default_f32("serve") == "f32c"
    args = types.SimpleNamespace(f32="f32r")
    assert bench.f32_forms(args) == ["f32r", "f32c"]
    args.f32 = "f32c"
    assert bench.f32_forms(args) == ["f32c", "f32r"]


def test_relative_error_helper():
    a = (np.array([1.0, 0.0]), np.array([0.0, 2.0]), np.array([0.0, 0.0]))
    ref = (np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.array([0.0, 0.0]))
    mx, med = bench.rel_err(a, ref)
    assert mx == 1.0 and med == 0.5
    json.dumps({"max": mx, "median": med})   # plain floats for the JSON line
