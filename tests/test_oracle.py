"""Pins the oracle (oracle/bf_oracle.c) before it is trusted as the checker.

Sources of truth, in order:
  * golden fixtures written by the compiled reference (tests/golden/*.json,
    produced by tests/golden/make_golden.py from /root/reference's own
    cpu.cxx, arraysizes.cxx, adler32.c and icomplex4.hxx);
  * the reference's hard-coded checksum table (arraysizes.cxx:54-59) and its
    static_asserts (icomplex4.hxx:62-73);
  * closed forms: brute-force complex arithmetic on tiny shapes, impulses,
    G = 0, ties at G = 2^-s, saturation, the x86 lrint edge semantics.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import bf_oracle as bo

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


# ---------------------------------------------------------------- codec
def test_codec_static_asserts():
    # icomplex4.hxx:62-73
    assert bo.enc(1, 2) == 0x9A and bo.enc(-1, 2) == 0x7A
    assert bo.enc(1, -2) == 0x96 and bo.enc(-1, -2) == 0x76
    assert bo.dec(0x9A) == (1, 2) and bo.dec(0x7A) == (-1, 2)
    assert bo.dec(0x96) == (1, -2) and bo.dec(0x76) == (-1, -2)


def test_codec_matches_reference_tables():
    g = load("codec.json")
    for x, (re, im) in enumerate(g["decode"]):
        assert bo.dec(x) == (re, im), x
    for re, im, byte in g["encode"]:
        assert bo.enc(re, im) == byte, (re, im)


def test_codec_offset_binary_closed_form():
    # each nibble stores value + 8: value = nibble - 8 for all 16 nibbles
    for x in range(256):
        assert bo.dec(x) == ((x >> 4) - 8, (x & 15) - 8)


# ---------------------------------------------------------------- adler32
def test_adler32_known_answers():
    assert bo.adler32(np.frombuffer(b"Wikipedia", np.uint8)) == 0x11E60398
    assert bo.adler32(np.zeros(0, np.uint8)) == 1
    assert bo.adler32(np.frombuffer(b"a", np.uint8)) == 0x00620062


def test_adler32_matches_reference():
    for case in load("adler32.json")["cases"]:
        buf = bo.fill_uniform(case["len"], case["seed"])
        assert bo.adler32(buf) == case["adler32"], case


# ---------------------------------------------------------------- generators
def test_splitmix64_known_answers():
    # SplitMix64 with state 0: published first outputs
    assert bo.splitmix64_at(0, 0) == 0xE220A8397B1DCDAF
    assert bo.splitmix64_at(0, 1) == 0x6E789E6AA1B965F4
    assert bo.splitmix64_at(0, 2) == 0x06C45D188009454F


def test_uniform_values_in_range_and_balanced():
    buf = bo.fill_uniform(1 << 16, 7)
    re = (buf >> 4).astype(int) - 8
    im = (buf & 15).astype(int) - 8
    assert re.min() == -7 and re.max() == 7 and im.min() == -7 and im.max() == 7
    for v in (re, im):
        counts = np.bincount(v + 7, minlength=15)
        assert counts.min() > 0.9 * len(v) / 15 and counts.max() < 1.1 * len(v) / 15
    # offset parameter = the same stream shifted
    assert np.array_equal(bo.fill_uniform(100, 7, offset=50), buf[50:150])


def test_uniform_mapping_formula():
    for i in range(64):
        z = bo.splitmix64_at(11, i)
        re = ((z & 0xFFFFFFFF) * 15 >> 32) - 7
        im = ((z >> 32) * 15 >> 32) - 7
        assert bo.fill_uniform(64, 11)[i] == bo.enc(re, im)


def test_phase_ramp_closed_forms():
    F, B, D = 2, 5, 200
    A = bo.phase_ramp(F, B, D).reshape(F, B, D)
    assert np.array_equal(A[0], A[1])
    for b in range(B):
        for d in range(D):
            k = (b * d) % 97
            re, im = bo.dec(int(A[0, b, d]))
            if k == 0:
                assert (re, im) == (7, 0)
            th = 2 * math.pi * k / 97
            assert re == round(7 * math.cos(th)) and im == round(7 * math.sin(th))
            # conjugate symmetry: k and 97 - k
    for k in range(1, 97):
        b, d = 1, k
        b2, d2 = 1, 97 - k
        assert bo.dec(int(A[0, b, d]))[0] == bo.dec(int(A[0, b2, d2]))[0]


def test_setup_gain_formula():
    # arraysizes.cxx:31-32 evaluated in float32, left to right
    nG, D = 96 * 4, 64
    G = bo.setup_gain(nG, D)
    for n in range(nG):
        g = np.float32(n) / np.float32(nG)
        g = np.float32(g / np.float32(D))
        g = np.float32(g * np.float32(15 + n % 15))
        g = np.float32(g / np.float32(30))
        assert G[n] == g


# ---------------------------------------------------------------- requant
def test_requant_ties_to_even_and_clamp():
    # G = 2^-s: G * raw = raw / 2^s exactly; lrint rounds ties to even
    assert [bo.requant(0.5, r) for r in (-5, -3, -1, 1, 3, 5)] == [-2, -2, 0, 0, 2, 2]
    assert [bo.requant(0.25, r) for r in (2, 6, 10, 14)] == [0, 2, 2, 4]
    assert bo.requant(1.0, 100) == 7 and bo.requant(1.0, -100) == -7
    assert bo.requant(0.0, 12345) == 0


def test_requant_x86_out_of_range_semantics():
    # int(lrint(y)) on x86-64: low 32 bits of the 64-bit result; NaN and
    # |y| >= 2^63 give INT64_MIN (low word 0)
    assert bo.requant(3.0e9, 1) == -7      # 3000000000 -> -1294967296 -> -7
    assert bo.requant(2.0 ** 32 + 2 ** 9, 1) == 7  # low word 512 -> 7
    assert bo.requant(1.0e30, 1) == 0
    assert bo.requant(float("inf"), 1) == 0
    assert bo.requant(float("nan"), 1) == 0


# ---------------------------------------------------------------- form_beams
def brute_force(E, A, G, T, F, D, B, P=2, jlayout=0):
    """Pure-Python complex arithmetic (tiny shapes only)."""
    def val(x):
        return complex((int(x) >> 4) - 8, (int(x) & 15) - 8)

    J = np.zeros(B * F * P * T, np.uint8)
    for b in range(B):
        for f in range(F):
            for p in range(P):
                for t in range(T):
                    acc = 0
                    for d in range(D):
                        acc += val(A[(f * B + b) * D + d]) * val(E[((t * F + f) * D + d) * P + p])
                    g = np.float32(G[f * B + b])
                    q = []
                    for part in (acc.real, acc.imag):
                        y = np.float32(g * np.float32(int(part)))
                        q.append(int(max(-7, min(7, np.rint(y)))))
                    off = ((b * F + f) * P + p) * T + t if jlayout == 0 else ((b * F + f) * T + t) * P + p
                    J[off] = (((q[0] + 8) << 4) | (q[1] + 8))
    return J


@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (3, 2, 5, 3), (8, 1, 32, 8), (5, 2, 9, 4)])
@pytest.mark.parametrize("jlayout", [0, 1])
def test_brute_force_tiny(shape, jlayout):
    T, F, D, B = shape
    E = bo.fill_uniform(T * F * D * 2, 3)
    A = bo.fill_uniform(F * B * D, 4)
    G = (np.arange(F * B, dtype=np.float32) + 1) / np.float32(8 * D)
    J = bo.form_beams(E, A, G, T, F, D, B, jlayout=jlayout)
    assert np.array_equal(J, brute_force(E, A, G, T, F, D, B, jlayout=jlayout))


# reference checksum table (arraysizes.cxx:54-59; the last two entries belong
# to the F = 1 configs of arraysizes.hxx:58-70) plus the shipped benchmark
# shape and the small workload shape (adler32 of J in Jlinear order)
REF_CHECKSUMS = [
    ((1, 1, 1, 1), 0x019A0111),
    ((8, 1, 32, 8), 0x3A3344BC),
    ((32, 32, 512, 128), 0xAB69DFEE),
    ((32, 32, 512, 96), 0x691BE8A7),
    ((256, 4, 64, 16), 0x07CEE085),
]


@pytest.mark.parametrize("shape,checksum", REF_CHECKSUMS)
def test_reference_checksums(shape, checksum):
    T, F, D, B = shape
    E, A, G = bo.setup_pattern(T, F, D, B)
    J = bo.form_beams(E, A, G, T, F, D, B)
    assert bo.adler32(J) == checksum


@pytest.mark.slow
@pytest.mark.parametrize("shape,checksum", [((32768, 32, 512, 96), 0xC4FB6A4C), ((32768, 16, 512, 96), 0x27303EA9)])
def test_reference_checksums_full_size(shape, checksum):
    T, F, D, B = shape
    E, A, G = bo.setup_pattern(T, F, D, B)
    J = bo.form_beams(E, A, G, T, F, D, B)
    assert bo.adler32(J) == checksum


def _golden_shape(g):
    s = g["shape"]
    return s["ntimes"], s["nfrequencies"], s["ndishes"], s["nbeams"]


def test_reference_pattern_at_compiled_shape():
    g = load("ref_pattern.json")
    T, F, D, B = _golden_shape(g)
    E, A, G = bo.setup_pattern(T, F, D, B)
    idx = np.array([s[0] for s in g["samples"]], np.uint64)
    want = np.array([s[1] for s in g["samples"]], np.uint8)
    assert np.array_equal(bo.form_beams_points(idx, E, A, G, T, F, D, B), want)
    J = bo.form_beams(E, A, G, T, F, D, B)
    assert bo.adler32(J) == g["adler32"]


@pytest.mark.parametrize("run", [0, 1])
def test_reference_seeded_at_compiled_shape(run):
    g = load("ref_seeded.json")["runs"][run]
    T, F, D, B = _golden_shape(g)
    seed = g["seed"]
    E = bo.fill_uniform(T * F * D * 2, seed)
    A = bo.fill_uniform(F * B * D, seed ^ 0xA5A5A5A5A5A5A5A5)
    G = bo.setup_gain(F * B, D)
    idx = np.array([s[0] for s in g["samples"]], np.uint64)
    want = np.array([s[1] for s in g["samples"]], np.uint8)
    assert np.array_equal(bo.form_beams_points(idx, E, A, G, T, F, D, B), want)
    if run == 0:
        J = bo.form_beams(E, A, G, T, F, D, B)
        assert bo.adler32(J) == g["adler32"]


def test_impulse_closed_form():
    # one nonzero A and E entry: J = requant(G A E) at one output, 0 elsewhere
    T, F, D, B = 9, 3, 16, 5
    E = np.full(T * F * D * 2, bo.enc(0, 0), np.uint8)
    A = np.full(F * B * D, bo.enc(0, 0), np.uint8)
    t, f, d, p, b = 4, 1, 7, 1, 3
    E[((t * F + f) * D + d) * 2 + p] = bo.enc(2, -3)
    A[(f * B + b) * D + d] = bo.enc(-1, 2)
    G = np.zeros(F * B, np.float32)
    G[f * B + b] = 0.5
    J = bo.form_beams(E, A, G, T, F, D, B)
    prod = complex(-1, 2) * complex(2, -3)  # (4 + 7j)
    want = np.full(J.size, bo.enc(0, 0), np.uint8)
    want[((b * F + f) * 2 + p) * T + t] = bo.enc(int(np.rint(np.float32(0.5 * prod.real))), int(np.rint(0.5 * prod.imag)))
    assert np.array_equal(J, want)


def test_gain_zero_gives_all_0x88():
    T, F, D, B = 16, 2, 32, 4
    E = bo.fill_uniform(T * F * D * 2, 1)
    A = bo.fill_uniform(F * B * D, 2)
    J = bo.form_beams(E, A, np.zeros(F * B, np.float32), T, F, D, B)
    assert (J == 0x88).all()


def test_layouts_are_permutations():
    T, F, D, B = 12, 3, 16, 5
    E = bo.fill_uniform(T * F * D * 2, 5)
    A = bo.fill_uniform(F * B * D, 6)
    G = bo.setup_gain(F * B, D) * np.float32(16)
    J1 = bo.form_beams(E, A, G, T, F, D, B, jlayout=0).reshape(B, F, 2, T)
    J2 = bo.form_beams(E, A, G, T, F, D, B, jlayout=1).reshape(B, F, T, 2)
    assert np.array_equal(J1.transpose(0, 1, 3, 2), J2)


def test_linearity_in_E_before_requant():
    # with G = 1 and tiny values, J is exact (no rounding, no clamp), so
    # J(E) for a single-dish shape equals A * E
    T, F, D, B = 4, 1, 1, 3
    for ar in (-2, 0, 1):
        for er in (-3, 0, 2):
            E = np.full(T * F * D * 2, bo.enc(er, 1), np.uint8)
            A = np.full(F * B * D, bo.enc(ar, -1), np.uint8)
            J = bo.form_beams(E, A, np.ones(F * B, np.float32), T, F, D, B)
            prod = complex(ar, -1) * complex(er, 1)
            assert (J == bo.enc(int(prod.real), int(prod.imag))).all()
