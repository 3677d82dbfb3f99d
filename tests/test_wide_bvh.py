"""The host construction of the fast traversal mode's trees, checked on the
CPU: tests/cpu/wide_bvh_check.cpp, built with the host compiler against
muon_amd/csrc/device/wide_bvh.h (the header the scene upload uses), runs the
SAH-optimal 4-wide collapse and the 8-bit quantizer on random binary trees
and checks the wide tree's structure (every primitive in one leaf, child
boxes containing their subtrees, the stack bound), its cost against the
greedy collapse and, on small trees, against an enumeration of every
collapse, and that every quantized plane lies outside the float box."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_wide_collapse_and_quantizer(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "wide_bvh_check"
    src = os.path.join(ROOT, "tests", "cpu", "wide_bvh_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "muon_amd", "csrc", "device"), src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
