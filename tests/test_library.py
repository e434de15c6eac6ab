"""CPU-side checks of libbeamform_amd.so: it loads, exports every function
declared in include/beamform.h, and its host-only paths and argument checks
behave (no kernel is launched here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import beamforming_amd as bf
from oracle import bf_oracle as bo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(bf_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", bf.library_path], check=True, capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(bf.library_path)
    for n in names:
        assert getattr(lib, n)


def test_status_strings_and_version():
    assert bf.lib.bf_status_string(0) == b"ok"
    assert bf.lib.bf_status_string(2) == b"unsupported shape"
    assert bf.lib.bf_status_string(999) == b"unknown status"
    assert b"gfx950" in bf.lib.bf_version()


def test_host_adler32_matches_oracle():
    for n in (0, 1, 15, 16, 5551, 5552, 5553, 100000):
        buf = bo.fill_uniform(n, n + 1)
        assert bf.adler32(buf) == bo.adler32(buf)
    assert bf.adler32(np.frombuffer(b"Wikipedia", np.uint8)) == 0x11E60398


@pytest.mark.parametrize("bad", [
    dict(ntimes=0), dict(nfrequencies=-1), dict(ndishes=0), dict(nbeams=0), dict(j_layout=7),
])
def test_invalid_shapes_rejected_before_any_device_call(bad):
    s = dict(ntimes=16, nfrequencies=1, ndishes=32, nbeams=16, npolarizations=2, j_layout=0)
    s.update(bad)
    shape = bf.Shape.make(s["ntimes"], s["nfrequencies"], s["ndishes"], s["nbeams"], s["npolarizations"],
                          s["j_layout"])
    A = np.zeros(max(1, abs(s["nfrequencies"] * s["nbeams"] * s["ndishes"])), np.uint8)
    G = np.zeros(max(1, abs(s["nfrequencies"] * s["nbeams"])), np.float32)
    h = ctypes.c_void_p()
    rc = bf.lib.bf_plan_create(ctypes.byref(shape), A.ctypes.data_as(ctypes.c_void_p),
                               G.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
    assert rc == 1  # BF_ERR_INVALID_ARGUMENT
    assert not h.value
    assert bf.lib.bf_last_error()


@pytest.mark.parametrize("extents", [
    (2**34 + 1, 1, 32, 16),      # longer than the time tile indices allow
    (2**30, 2**20, 65536, 16),   # E bytes beyond 2^62
    (2**33, 2**20, 32, 2**20),   # J bytes beyond 2^62
])
def test_oversized_shapes_rejected_before_any_device_call(extents):
    shape = bf.Shape.make(*extents)
    A = np.zeros(16, np.uint8)
    G = np.zeros(16, np.float32)
    h = ctypes.c_void_p()
    rc = bf.lib.bf_plan_create(ctypes.byref(shape), A.ctypes.data_as(ctypes.c_void_p),
                               G.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h))
    assert rc == 2  # BF_ERR_UNSUPPORTED
    assert not h.value
    assert b"too large" in bf.lib.bf_last_error()


def test_null_arguments_rejected():
    shape = bf.Shape.make(16, 1, 32, 16)
    h = ctypes.c_void_p()
    assert bf.lib.bf_plan_create(ctypes.byref(shape), None, None, ctypes.byref(h)) == 1
    assert bf.lib.bf_form_beams_device(None, None, None, None) == 1
    assert bf.lib.bf_plan_destroy(None) == 0
    with pytest.raises(bf.BeamformError):
        bf.adler32_device(0, 16)


def test_no_device_is_reported_not_fatal():
    # This container has no GPU: a valid plan request must fail cleanly.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    shape = bf.Shape.make(16, 1, 32, 16)
    A = np.zeros(shape.a_bytes, np.uint8)
    G = np.zeros(shape.g_count, np.float32)
    with pytest.raises(bf.BeamformError) as ei:
        bf.Plan(shape, A, G)
    assert ei.value.status in (3, 8)  # device error / no device


@pytest.mark.parametrize("bad", [
    dict(nfrequencies=8, ndishes=64, chunk_ntimes=16, group_channels=3, ring_slots=2),   # G does not divide F
    dict(nfrequencies=8, ndishes=64, chunk_ntimes=16, group_channels=4, ring_slots=1),   # one slot
    dict(nfrequencies=8, ndishes=64, chunk_ntimes=0, group_channels=4, ring_slots=2),    # empty chunks
    dict(nfrequencies=8, ndishes=64, chunk_ntimes=16, group_channels=4, ring_slots=2, npolarizations=1),
])
def test_ingest_config_rejected_before_any_device_call(bad):
    args = dict(bad)
    p = args.pop("npolarizations", 2)
    cfg = bf.IngestConfig(args["nfrequencies"], args["ndishes"], p, args["chunk_ntimes"], args["group_channels"],
                          args["ring_slots"], 0, 0)
    h = ctypes.c_void_p()
    assert bf.lib.bf_ingest_create(ctypes.byref(cfg), ctypes.byref(h)) == bf.BF_ERR_INVALID_ARGUMENT
    assert not h.value
