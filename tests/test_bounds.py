"""Host check of the gfx950 kernels' memory accesses (tests/bounds_check.cpp):
the tile pipeline's address arithmetic, shared with the kernels through
bf_device.hpp, run for every wave and lane of plain and batched launches over
random shapes and chunk lists. Every E load and J store must stay inside its
buffer, every J byte must be written exactly once, and every J byte that goes
through the LDS staging area must come from the staging byte holding its
time. Compiled for the CPU with hipcc (no GPU)."""
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
def test_kernel_accesses_stay_in_bounds(tmp_path):
    exe = tmp_path / "bounds_check"
    cmd = [HIPCC, "-std=c++17", "-O2", "-Wall", "-I", CSRC, "-I", INCLUDE, os.path.join(HERE, "bounds_check.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_are_caught(tmp_path, plant):
    # the model's own planted mistakes: no row clamp at T - 1, a read of the
    # other staging slot, staging writes one row off
    exe = tmp_path / f"bounds_check_plant{plant}"
    cmd = [HIPCC, "-std=c++17", "-O2", f"-DPLANT={plant}", "-I", CSRC, "-I", INCLUDE,
           os.path.join(HERE, "bounds_check.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode != 0 and "FAIL" in run.stdout, run.stdout[-2000:]
