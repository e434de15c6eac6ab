"""Host checks of the serving loop's admission policy and arena allocator
(bf_sched.hpp), compiled with g++ (no GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "beamforming_amd", "csrc")


def test_sched_policy(tmp_path):
    exe = tmp_path / "sched_check"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined",
           "-I", CSRC, os.path.join(HERE, "sched_check.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "sanitize" in build.stderr:
        cmd = [c for c in cmd if not c.startswith("-fsanitize")]
        build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")
