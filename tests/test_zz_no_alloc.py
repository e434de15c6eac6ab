"""GPU test of bf_adler32_device_noalloc, the allocation-free variant of
bf_adler32_device (static device scratch), called through the C ABI with
ctypes: two calls in a row, the second on a smaller input, each equal to
bf_adler32_device on the same bytes and to the oracle's Adler-32."""
import ctypes

import numpy as np
import pytest
import torch

import beamforming_amd as bf
from oracle import bf_oracle as bo

pytestmark = pytest.mark.gpu


def c_abi():
    lib = ctypes.CDLL(bf.library_path)
    for name in ("bf_adler32_device_noalloc", "bf_adler32_device"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p]
    return lib


def call(fn, ptr, n, stream):
    out = ctypes.c_uint32(0xFFFFFFFF)
    rc = fn(ctypes.c_void_p(ptr), n, ctypes.byref(out), ctypes.c_void_p(stream))
    assert rc == 0, f"status {rc}"
    return int(out.value)


def test_noalloc_adler32_matches_the_allocating_entry_point():
    lib = c_abi()
    stream = torch.cuda.current_stream().cuda_stream
    big = (64 << 20) + 13  # the partial pass at its 2048 workgroups
    buf = torch.randint(0, 256, (big + 64,), dtype=torch.uint8, device="cuda")
    host = buf.cpu().numpy()
    # two calls in a row: the first fills every partial slot of the scratch,
    # the second (smaller, 25 workgroups) must not see the leftovers
    for n, off in [(big, 3), (100003, 5), (0, 0), (17, 1)]:
        ptr = buf.data_ptr() + off
        got = call(lib.bf_adler32_device_noalloc, ptr, n, stream)
        ref = call(lib.bf_adler32_device, ptr, n, stream)
        assert got == ref, f"n={n}: noalloc 0x{got:08x}, bf_adler32_device 0x{ref:08x}"
        assert got == bo.adler32(host[off:off + n]), f"n={n}"


def test_noalloc_rejects_null_arguments():
    lib = c_abi()
    out = ctypes.c_uint32()
    assert lib.bf_adler32_device_noalloc(None, 16, ctypes.byref(out), None) == bf.BF_ERR_INVALID_ARGUMENT
