"""The request server's threading on the CPU, under ThreadSanitizer.

n-body-sim_amd/csrc/server.cpp is built with clang -fsanitize=thread against
a test-only CPU stand-in for the engine (tests/server_host/engine_stub.cpp:
host arrays and a free drift instead of device kernels), and
tests/server_host/server_test.cpp drives it:
  * the admission pass on built inputs (placement replayed on a copy of the
    free ranges, bypasses counted only when a later request is admitted, the
    starvation limit, the slot limit);
  * a cancel that arrives while the request is being uploaded is kept;
  * submit / poll / cancel from several threads at once with snapshot
    backpressure, every request's events checked (order, snapshot steps,
    final state bit for bit, cancel semantics).
The run must pass its checks with no ThreadSanitizer report.  (GCC 11's
ThreadSanitizer does not intercept pthread_cond_clockwait, which libstdc++'s
timed waits use, and reports false races on them; ROCm's clang runtime
does.)
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "server_host")
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def server_binary(tmp_path_factory):
    if not os.path.exists(CLANG) or shutil.which("ld") is None:
        pytest.skip("clang++ with ThreadSanitizer is not available")
    out = str(tmp_path_factory.mktemp("server_host") / "server_tsan")
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=thread",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "n-body-sim_amd", "csrc", "server.cpp"),
           os.path.join(SRC, "engine_stub.cpp"), os.path.join(SRC, "server_test.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("part", ["admission", "cancel", "stress"])
def test_server_under_thread_sanitizer(server_binary, part):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    r = subprocess.run([server_binary, part], capture_output=True, text=True,
                       timeout=300, env=env)
    text = r.stdout + r.stderr
    assert "ThreadSanitizer" not in text, text[-4000:]
    assert r.returncode == 0 and text.rstrip().endswith("ok"), text[-4000:]
