"""Host checks of the kernels' per-lane arithmetic (bf_device.hpp: nibble
unpack, operand windows, weight packing, nibble-offset corrections, fast and
exact requantisation, output packing) against the oracle, compiled for the
CPU with hipcc (no GPU)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..")
CSRC = os.path.join(ROOT, "beamforming_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_device_math(tmp_path):
    subprocess.run(["make", "-s", "-C", ORACLE], check=True)
    exe = tmp_path / "device_math_check"
    # -ffp-contract=off as in the library build: the gain multiply stays a
    # separately rounded operation
    cmd = [HIPCC, "-std=c++17", "-O1", "-ffp-contract=off", "-Wall", "-I", CSRC, "-I", ORACLE,
           os.path.join(HERE, "device_math_check.cpp"), "-L", ORACLE, "-lbf_oracle", f"-Wl,-rpath,{ORACLE}",
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("ok")
