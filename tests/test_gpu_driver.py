"""The reference's driver flow end to end on the GPU: the C++ harness
beamforming_amd/bf_form_beams (setup() inputs -> form_beams on host arrays
through include/beamform.hpp -> adler32 check, as in cpu.cxx:48-67) must
reproduce the reference's checksums (arraysizes.cxx:54-59 and the table in
the harness, computed with the reference cpu.cxx). For the J2 layout the
expected checksum comes from the oracle on the same inputs."""
import os
import subprocess
import sys

import pytest

from oracle import bf_oracle as bo

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "beamforming_amd", "bf_form_beams")


def run(args):
    print(f"bf_form_beams {' '.join(args)}", flush=True)
    sys.stdout.flush()
    assert os.path.exists(EXE), "bf_form_beams is built with the library (make -C beamforming_amd)"
    out = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=300)
    return out.returncode, out.stdout + out.stderr


@pytest.mark.parametrize("dims", [
    (1, 1, 1, 1),            # reference table (arraysizes.cxx:59)
    (8, 1, 32, 8),           # reference table (arraysizes.cxx:58)
    (32, 32, 512, 96),
    (256, 4, 64, 16),
    (32768, 32, 512, 96),    # the full-size workload (0xc4fb6a4c)
])
def test_reference_checksums(dims):
    rc, out = run([str(d) for d in dims])
    assert rc == 0, out
    assert "Checksum matches." in out, out


def test_j2_layout_checksum_from_oracle():
    T, F, D, B = 256, 4, 64, 16
    E, A, G = bo.setup_pattern(T, F, D, B)
    want = bo.adler32(bo.form_beams(E, A, G, T, F, D, B, jlayout=1))
    rc, out = run([str(T), str(F), str(D), str(B), "--layout", "bftp", "--expect", f"0x{want:08x}"])
    assert rc == 0, out
    assert "Checksum matches." in out, out


def test_checksum_mismatch_is_reported():
    rc, out = run(["256", "4", "64", "16", "--expect", "0x12345678"])
    assert rc == 1 and "CHECKSUM MISMATCH" in out, out
