"""ctypes binding of the oracle (oracle/libbf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product library never uses
it. See bf_oracle.h for what each function restates and what pins it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbf_oracle.so")

BFO_J_BFPT = 0
BFO_J_BFTP = 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    lib.bfo_enc.restype = ctypes.c_uint8
    lib.bfo_enc.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.bfo_dec_re.restype = ctypes.c_int
    lib.bfo_dec_re.argtypes = [ctypes.c_uint8]
    lib.bfo_dec_im.restype = ctypes.c_int
    lib.bfo_dec_im.argtypes = [ctypes.c_uint8]
    lib.bfo_adler32.restype = ctypes.c_uint32
    lib.bfo_adler32.argtypes = [vp, sz]
    lib.bfo_setup_pattern.restype = None
    lib.bfo_setup_pattern.argtypes = [vp, sz, vp, sz, vp, sz, sz]
    lib.bfo_setup_gain.restype = None
    lib.bfo_setup_gain.argtypes = [vp, sz, sz]
    lib.bfo_fill_uniform.restype = None
    lib.bfo_fill_uniform.argtypes = [vp, sz, ctypes.c_uint64, ctypes.c_uint64]
    lib.bfo_splitmix64_at.restype = ctypes.c_uint64
    lib.bfo_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.bfo_fill_phase_ramp.restype = None
    lib.bfo_fill_phase_ramp.argtypes = [vp, sz, sz, sz]
    lib.bfo_form_beams.restype = ctypes.c_int
    lib.bfo_form_beams.argtypes = [vp, vp, vp, vp, sz, sz, sz, sz, sz, ctypes.c_int, ctypes.c_int]
    lib.bfo_form_beams_points.restype = ctypes.c_int
    lib.bfo_form_beams_points.argtypes = [vp, vp, sz, vp, vp, vp, sz, sz, sz, sz, sz, ctypes.c_int]
    lib.bfo_requant.restype = ctypes.c_int
    lib.bfo_requant.argtypes = [ctypes.c_float, ctypes.c_int]
    return lib


lib = _load()


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def enc(re, im):
    return int(lib.bfo_enc(re, im))


def dec(x):
    return int(lib.bfo_dec_re(x)), int(lib.bfo_dec_im(x))


def adler32(data):
    a = np.ascontiguousarray(data, dtype=np.uint8)
    return int(lib.bfo_adler32(_p(a), a.size))


def setup_pattern(T, F, D, B, P=2):
    """Reference setup() inputs (arraysizes.cxx:8-36) for runtime shapes."""
    E = np.empty(T * F * D * P, np.uint8)
    A = np.empty(F * B * D, np.uint8)
    G = np.empty(F * B, np.float32)
    lib.bfo_setup_pattern(_p(E), E.size, _p(A), A.size, _p(G), G.size, D)
    return E, A, G


def setup_gain(nG, D):
    G = np.empty(nG, np.float32)
    lib.bfo_setup_gain(_p(G), nG, D)
    return G


def fill_uniform(n, seed, offset=0):
    buf = np.empty(n, np.uint8)
    lib.bfo_fill_uniform(_p(buf), n, seed, offset)
    return buf


def splitmix64_at(seed, counter):
    return int(lib.bfo_splitmix64_at(seed, counter))


def phase_ramp(F, B, D):
    A = np.empty(F * B * D, np.uint8)
    lib.bfo_fill_phase_ramp(_p(A), F, B, D)
    return A


def form_beams(E, A, G, T, F, D, B, P=2, jlayout=BFO_J_BFPT, nthreads=0):
    E = np.ascontiguousarray(E, np.uint8)
    A = np.ascontiguousarray(A, np.uint8)
    G = np.ascontiguousarray(G, np.float32)
    assert E.size == T * F * D * P and A.size == F * B * D and G.size == F * B
    J = np.empty(B * F * P * T, np.uint8)
    rc = lib.bfo_form_beams(_p(J), _p(E), _p(A), _p(G), T, F, D, P, B, jlayout, nthreads)
    if rc != 0:
        raise ValueError("bfo_form_beams failed")
    return J


def form_beams_points(idx, E, A, G, T, F, D, B, P=2, jlayout=BFO_J_BFPT):
    idx = np.ascontiguousarray(idx, np.uint64)
    E = np.ascontiguousarray(E, np.uint8)
    A = np.ascontiguousarray(A, np.uint8)
    G = np.ascontiguousarray(G, np.float32)
    out = np.empty(idx.size, np.uint8)
    rc = lib.bfo_form_beams_points(_p(out), _p(idx), idx.size, _p(E), _p(A), _p(G), T, F, D, P, B, jlayout)
    if rc != 0:
        raise ValueError("bfo_form_beams_points failed")
    return out


def requant(gain, raw):
    return int(lib.bfo_requant(gain, raw))
