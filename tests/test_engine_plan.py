"""The engine's index arithmetic and the scheduler's policies, checked on
the CPU: tests/cpu/engine_plan_check.cpp, built with the host C++ compiler
against muon_amd/csrc/device/engine_plan.h and sched_policy.h (the headers
the kernels and the scheduler use), runs randomized checks of the buffer
layout, the admission window / plan / fill mapping of samples to jobs,
pixels and queue positions, the film range allocator, the slice choice's
starvation bound and the job-slot runs the admission flush copies."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_engine_index_math_and_policies(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "engine_plan_check"
    src = os.path.join(ROOT, "tests", "cpu", "engine_plan_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "muon_amd", "csrc", "device"), src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
