"""CPU checks of bench.py's contract and helpers (no GPU)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_workload_figures():
    # full-size single shot: 8 T F B D P int8 ops, E + A + G + J bytes
    assert bench.workload_ops(32768, 32, 512, 96) == 824_633_720_832
    assert bench.workload_bytes(32768, 32, 512, 96) == 32768 * 32 * 1024 + 32 * 96 * 512 + 4 * 32 * 96 + 96 * 32 * 2 * 32768


def test_amd_smi_metric_parsing():
    text = """GPU: 0
    POWER:
        SOCKET_POWER: 1401 W
        GFX_VOLTAGE: N/A
    CLOCK:
        GFX_0:
            CLK: 1613 MHz
            MIN_CLK: 500 MHz
        GFX_1:
            CLK: 1535 MHz
            MIN_CLK: 500 MHz
        MEM_0:
            CLK: 2000 MHz
"""
    power, clocks = bench.parse_amd_smi_metric(text)
    assert power == 1401.0
    assert clocks == [1613.0, 1535.0]
    assert bench.parse_amd_smi_metric("no reading") == (None, [])


def test_synthetic_inputs_match_the_oracle_generators():
    # bench.py builds its inputs without the oracle; they are the same bytes
    from oracle import bf_oracle as bo
    import numpy as np
    for n, D in [(32 * 96, 512), (64 * 96, 512), (16 * 512, 512), (7, 3)]:
        assert np.array_equal(bench.reference_gains(n, D), bo.setup_gain(n, D))
    assert np.array_equal(bench.phase_ramp_weights(16, 512, 512), bo.phase_ramp(16, 512, 512))


def test_roofline_picks_the_binding_bound():
    # full size: HBM binds (1.28 GB at 8 TB/s = 160 us against 82 us of fp6 MFMA)
    r = bench.roofline(bench.workload_ops(32768, 32, 512, 96), bench.workload_bytes(32768, 32, 512, 96), 0.3192)
    assert r["bound"] == "hbm" and abs(r["frac"] - 0.5) < 0.001 and r["frac"] == r["frac_hbm"]
    # wide beams: the fp6 MFMA bound binds (550 G ops at 10.1 POP/s = 54 us against 34 us of HBM)
    r = bench.roofline(bench.workload_ops(8192, 16, 512, 512), bench.workload_bytes(8192, 16, 512, 512), 0.1)
    assert r["bound"] == "mfma_fp6" and r["frac"] == r["frac_mfma_fp6"] and r["frac_mfma_int8"] > r["frac"]
