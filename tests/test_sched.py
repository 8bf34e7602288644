"""CPU: the search server's continuous-batching scheduler
(pydiskann_amd/csrc/spf_sched.h) against a simulated executor
(tests/sched_sim.cpp): per-row results independent of batching, simulated
device memory within the budget, request merging and concurrency, rejection
above the budget, error propagation, timeouts and drain on shutdown; once
plain and once under the thread sanitizer.  The sanitizer build uses ROCm's
clang: GCC 11's libtsan does not intercept pthread_cond_clockwait, which
libstdc++ uses for timed waits, and reports false races."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "sched_sim.cpp")
INC = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "pydiskann_amd", "csrc")]
CLANG = "/opt/rocm/llvm/bin/clang++"


def build_and_run(cxx, flags, tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-pthread", *flags, *INC, SRC, "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "sched ok" in r.stdout
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]


def test_scheduler_simulated(tmp_path):
    build_and_run(shutil.which("g++") or "g++", ["-Wall", "-Werror"], tmp_path, "sim")


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang not installed")
def test_scheduler_thread_sanitizer(tmp_path):
    build_and_run(CLANG, ["-fsanitize=thread"], tmp_path, "sim_tsan")
