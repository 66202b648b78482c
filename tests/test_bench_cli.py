"""CPU checks of bench.py's contract helpers (no GPU work)."""
import json
import os
import subprocess
import sys

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
    kt = {"force_ms": 1000.0, "force_launches": 10, "combine_ms": 20.0,
          "combine_launches": 10}
    # 3e12 interactions over 1 s of force launches: 60 TFLOP/s at 20 flops each
    r = bench.roofline("f32", 3.0e12, kt, step_rate=2.9e12, traffic=123)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s"
    assert abs(r["achieved"] - 60.0) < 1e-9
    assert r["peak"] == 157.3 and abs(r["frac"] - 60.0 / 157.3) < 1e-12
    assert r["traffic"] == 123
    assert r["force_kernel_avg_ms"] == 100.0 and r["combine_kernel_avg_ms"] == 2.0
    assert abs(r["frac_of_whole_step"] - 58.0 / 157.3) < 1e-12
    assert bench.roofline("f64", 1.0e12, kt)["peak"] == 78.6


def test_cpu_baseline_threads_follow_the_affinity_mask():
    threads, aff, logical = bench.host_cpus()
    assert aff == len(os.sched_getaffinity(0))
    assert threads == min(aff, bench.HOST_CPU_SHARE) and threads >= 1
    assert logical == os.cpu_count()


def test_measured_traffic_reads_the_committed_profile():
    t = bench.measured_traffic()
    assert set(t) == {"f32r", "f32c", "f64"}
    # tens of MB per launch at N=65536 (partial sums of the split range)
    assert all(1e7 < v < 1e9 for v in t.values())


def test_f32_forms():
    assert bench.other_f32("f32r") == "f32c" and bench.other_f32("f32c") == "f32r"
    assert set(bench.PRECS) == {"f32r", "f32c", "f64"}


def test_serving_stream_is_seeded():
    a = bench.serve_requests(6, "plummer,star")
    b = bench.serve_requests(6, "plummer,star")
    assert [len(q["b"]) for q in a] == [len(q["b"]) for q in b]
    assert all(32 <= len(q["b"]) <= 16384 and 10 <= q["steps"] <= 1000 for q in a)
    assert all(q["dt"] in (1e-3, 5e-4) for q in a)


def test_relative_error_helper():
    a = (np.array([1.0, 0.0]), np.array([0.0, 2.0]), np.array([0.0, 0.0]))
    ref = (np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.array([0.0, 0.0]))
    mx, med = bench.rel_err(a, ref)
    assert mx == 1.0 and med == 0.5
    json.dumps({"max": mx, "median": med})   # plain floats for the JSON line
