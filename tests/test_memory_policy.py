"""The plain form's device-memory policy (claysampler_amd/csrc/hip/
memory_policy.h) on the CPU: scratch sizes by render window against a
brute-force count, the longest window that fits a budget against a scan of
every window, the budget formula, and the allocation-failure chain against an
allocator with a hard limit (tests/memory_check.cpp). The GPU side of the
same chain is tests/test_gpu_memory.py."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_memory_policy(tmp_path):
    exe = str(tmp_path / "memory_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "memory_check.cpp")], check=True)
    for seed in (1, 2, 3):
        out = subprocess.run([exe, str(seed), "300"], check=True, capture_output=True, text=True).stdout
        assert out.strip().endswith("failures 0"), out[-2000:]
