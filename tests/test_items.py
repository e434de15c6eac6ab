"""Host checks of the batched launch's work items (bf_items.hpp: balanced
and fixed-size items over random chunk lists), compiled for the CPU with
hipcc (no GPU)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..")
CSRC = os.path.join(ROOT, "beamforming_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_work_items(tmp_path):
    exe = tmp_path / "items_check"
    cmd = [HIPCC, "-std=c++17", "-O1", "-Wall", "-I", CSRC, "-I", INCLUDE, os.path.join(HERE, "items_check.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")
