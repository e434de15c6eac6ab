"""GPU parity: the HIP path (through the C ABI) against the oracle, bit-exact.

Every J buffer is prefilled with the sentinel byte 0x00, which decodes to
(-8, -8) and can never be produced (outputs are clamped to +-7), so "no 0x00
left" proves that every output was written.
"""
import sys

import numpy as np
import pytest
import torch

import beamforming_amd as bf
from oracle import bf_oracle as bo

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stream():
    return torch.cuda.current_stream().cuda_stream


def run_device(shape, E, A, G, plan=None):
    own = plan is None
    if own:
        plan = bf.Plan(shape, A, G)
    try:
        dE = torch.from_numpy(np.ascontiguousarray(E)).cuda()
        dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
        return dJ.cpu().numpy(), plan.describe()
    finally:
        if own:
            plan.close()


def seeded(T, F, D, B, seed):
    E = bo.fill_uniform(T * F * D * 2, seed)
    A = bo.fill_uniform(F * B * D, seed ^ 0xA5A5A5A5A5A5A5A5)
    G = bo.setup_gain(F * B, D)
    return E, A, G


def check_full(shape, E, A, G, want=None):
    T, F, D, B = shape.ntimes, shape.nfrequencies, shape.ndishes, shape.nbeams
    say(f"case T={T} F={F} D={D} B={B} P={shape.npolarizations} layout={shape.j_layout}")
    J, desc = run_device(shape, E, A, G)
    say(f"  kernel: {desc}")
    if want is None:
        want = bo.form_beams(E, A, G, T, F, D, B, P=shape.npolarizations, jlayout=shape.j_layout)
    assert (J != 0).all(), "unwritten outputs"
    bad = np.flatnonzero(J != want)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:8]}: got {J[bad[:8]]} want {want[bad[:8]]}"
    return J, desc


# ---------------------------------------------------------------- workloads
@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_small_config_seeded(layout):
    # T=256 F=4 D=64 P=2 B=16, uniform [-7, 7]
    shape = bf.Shape.make(256, 4, 64, 16, j_layout=layout)
    E, A, G = seeded(256, 4, 64, 16, 1)
    _, desc = check_full(shape, E, A, G)
    assert desc.startswith("mfma")


REF_CHECKSUMS = [
    ((1, 1, 1, 1), 0x019A0111),
    ((8, 1, 32, 8), 0x3A3344BC),
    ((32, 32, 512, 128), 0xAB69DFEE),
    ((32, 32, 512, 96), 0x691BE8A7),
    ((256, 4, 64, 16), 0x07CEE085),
]


@pytest.mark.parametrize("dims,checksum", REF_CHECKSUMS)
def test_reference_checksums(dims, checksum):
    T, F, D, B = dims
    shape = bf.Shape.make(T, F, D, B)
    E, A, G = bo.setup_pattern(T, F, D, B)
    J, _ = check_full(shape, E, A, G)
    assert bf.adler32(J) == checksum


def test_full_size_pattern_checksum():
    # 32768 x 32 x 512 x 2 x 96 with the reference setup() inputs:
    # arraysizes.cxx:56 gives 0xc4fb6a4c
    T, F, D, B = 32768, 32, 512, 96
    shape = bf.Shape.make(T, F, D, B)
    say("case full-size pattern")
    E, A, G = bo.setup_pattern(T, F, D, B)
    with bf.Plan(shape, A, G) as plan:
        say(f"  kernel: {plan.describe()}")
        dE = torch.from_numpy(E).cuda()
        del E
        dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
        assert bool((dJ != 0).all())
        assert bf.adler32_device(dJ.data_ptr(), shape.j_bytes, stream()) == 0xC4FB6A4C
        assert bf.adler32(dJ.cpu().numpy()) == 0xC4FB6A4C


def sampled_check(shape, dE, A, G, dJ, npoints, seed):
    T, F, D, B = shape.ntimes, shape.nfrequencies, shape.ndishes, shape.nbeams
    E = dE.cpu().numpy()
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, shape.j_bytes, npoints, dtype=np.uint64)
    # plus the corners of the index space
    idx = np.concatenate([idx, np.array([0, shape.j_bytes - 1], np.uint64)])
    want = bo.form_beams_points(idx, E, A, G, T, F, D, B, jlayout=shape.j_layout)
    got = dJ.cpu().numpy()[idx.astype(np.int64)]
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches of {idx.size}"


@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_full_size_seeded_sampled(layout):
    T, F, D, B = 32768, 32, 512, 96
    shape = bf.Shape.make(T, F, D, B, j_layout=layout)
    say(f"case full-size seeded layout={layout}")
    A = bo.fill_uniform(F * B * D, 12)
    G = bo.setup_gain(F * B, D)
    dE = torch.empty(shape.e_bytes, dtype=torch.uint8, device="cuda")
    bf.fill_uniform_device(dE.data_ptr(), shape.e_bytes, 11, 0, stream())
    dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
    with bf.Plan(shape, A, G) as plan:
        say(f"  kernel: {plan.describe()}")
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
    assert bool((dJ != 0).all())
    sampled_check(shape, dE, A, G, dJ, 50000, 1)


def test_wide_beam_phase_ramp_sampled():
    # T=8192 F=16 D=512 P=2 B=512, A = phase ramp, E uniform
    T, F, D, B = 8192, 16, 512, 512
    shape = bf.Shape.make(T, F, D, B)
    say("case wide-beam phase ramp")
    A = bo.phase_ramp(F, B, D)
    G = bo.setup_gain(F * B, D) * np.float32(4)
    dE = torch.empty(shape.e_bytes, dtype=torch.uint8, device="cuda")
    bf.fill_uniform_device(dE.data_ptr(), shape.e_bytes, 21, 0, stream())
    dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
    with bf.Plan(shape, A, G) as plan:
        say(f"  kernel: {plan.describe()}")
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
    assert bool((dJ != 0).all())
    sampled_check(shape, dE, A, G, dJ, 50000, 2)


# ---------------------------------------------------------------- edges
RAGGED = [
    (1, 1, 8, 1), (3, 2, 8, 5), (17, 1, 24, 17), (100, 3, 40, 33), (1001, 2, 64, 100), (64, 1, 520, 16),
    (33, 2, 1024, 20), (250, 1, 200, 200), (16, 5, 32, 96), (4100, 1, 512, 96),
]


@pytest.mark.parametrize("dims", RAGGED)
@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_ragged_shapes(dims, layout):
    T, F, D, B = dims
    shape = bf.Shape.make(T, F, D, B, j_layout=layout)
    E, A, G = seeded(T, F, D, B, T + 7 * D + B)
    G = G * np.float32(8)
    check_full(shape, E, A, G)


def random_shapes(n, seed, budget=2e8):
    """n shapes on the MFMA path (D a multiple of 8 up to 1024), T F B D at
    most `budget`, with a J layout and a gain scale each."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        T, F = int(rng.integers(1, 3000)), int(rng.integers(1, 5))
        D, B = 8 * int(rng.integers(1, 129)), int(rng.integers(1, 301))
        if T * F * B * D > budget:
            continue
        out.append(((T, F, D, B), int(rng.integers(0, 2)), float(rng.choice([1.0, 8.0, 64.0]))))
    return out


@pytest.mark.parametrize("dims,layout,scale", random_shapes(24, 2026))
def test_random_shapes(dims, layout, scale):
    T, F, D, B = dims
    shape = bf.Shape.make(T, F, D, B, j_layout=layout)
    E, A, G = seeded(T, F, D, B, T * 131 + D * 7 + B)
    check_full(shape, E, A, G * np.float32(scale))


@pytest.mark.parametrize("dims,P", [((7, 2, 5, 3), 2), ((9, 1, 16, 4), 1), ((5, 2, 12, 6), 3), ((6, 1, 1100, 3), 2)])
def test_generic_kernel_shapes(dims, P):
    T, F, D, B = dims
    shape = bf.Shape.make(T, F, D, B, npolarizations=P)
    E = bo.fill_uniform(T * F * D * P, 5)
    A = bo.fill_uniform(F * B * D, 6)
    G = bo.setup_gain(F * B, D) * np.float32(8)
    _, desc = check_full(shape, E, A, G)
    assert desc.startswith("generic")


def test_out_of_range_gains_follow_x86_semantics():
    T, F, D, B = 40, 2, 64, 8
    shape = bf.Shape.make(T, F, D, B)
    E, A, _ = seeded(T, F, D, B, 9)
    G = np.array([3.0e9, 1.0e30, np.inf, -np.inf, np.nan, 4.3e9, -2.2e9, 0.5] * 2, np.float32)
    _, desc = check_full(shape, E, A, G)
    assert desc.startswith("generic")


def test_gain_zero_ties_and_saturation():
    T, F, D, B = 64, 2, 64, 32
    shape = bf.Shape.make(T, F, D, B)
    E, A, _ = seeded(T, F, D, B, 10)
    G = np.zeros(F * B, np.float32)
    J, desc = check_full(shape, E, A, G)
    assert desc.startswith("mfma") and (J == 0x88).all()
    G = np.array([2.0 ** -s for s in range(F * B)], np.float32)  # exact halves -> ties to even
    check_full(shape, E, A, G)
    G = np.full(F * B, 3.0, np.float32)  # mostly saturated
    check_full(shape, E, A, G)
    G = np.where(np.arange(F * B) % 2 == 0, -1.0, 1.0).astype(np.float32) / np.float32(3)
    check_full(shape, E, A, G)


def test_nibble_minus_eight_inputs():
    # bytes with a 0 nibble decode to -8 (icomplex4.hxx:36-37)
    T, F, D, B = 48, 1, 64, 16
    shape = bf.Shape.make(T, F, D, B)
    rng = np.random.default_rng(3)
    E = rng.integers(0, 256, T * F * D * 2, dtype=np.uint8)
    A = rng.integers(0, 256, F * B * D, dtype=np.uint8)
    G = np.full(F * B, 1.0 / 64, np.float32)
    check_full(shape, E, A, G)


@pytest.mark.parametrize("D", [512, 200, 64])
@pytest.mark.parametrize("a_byte,e_kind", [(0x00, "random"), (0xFF, "random"), (0x0F, "random"), (0xF0, "random"),
                                           (0x00, 0x00), (0xFF, 0x00), (0x00, 0xFF), (0x0F, 0xF0)])
def test_extreme_operands(D, a_byte, e_kind):
    # The MFMA operands at their extremes: weights -8 and 7 in every field
    # position of the fp6 operands (E3M2 codes with and without sign), E
    # nibbles 0 and 15 in every position (E2M3 codes 0 and 15 after the
    # complement of one variant), at 64 and 512 dishes and with a dish tail
    # (200). The raw sums reach 128 D in magnitude; the gains keep them in
    # the requantisation's range and away from saturation.
    T, F, B = 96, 2, 40
    shape = bf.Shape.make(T, F, D, B)
    rng = np.random.default_rng(D + a_byte)
    if e_kind == "random":
        E = rng.integers(0, 256, T * F * D * 2, dtype=np.uint8)
    else:
        E = np.full(T * F * D * 2, e_kind, np.uint8)
    A = np.full(F * B * D, a_byte, np.uint8)
    A[rng.integers(0, A.size, A.size // 7)] = rng.integers(0, 256, A.size // 7, dtype=np.uint8)
    G = (rng.uniform(0.5, 7.0, F * B) / (16.0 * D)).astype(np.float32)
    check_full(shape, E, A, G)


def test_impulse():
    T, F, D, B = 32, 3, 64, 16
    shape = bf.Shape.make(T, F, D, B)
    E = np.full(T * F * D * 2, 0x88, np.uint8)
    A = np.full(F * B * D, 0x88, np.uint8)
    t, f, d, p, b = 17, 2, 45, 1, 9
    E[((t * F + f) * D + d) * 2 + p] = bo.enc(2, -2)
    A[(f * B + b) * D + d] = bo.enc(-1, 1)
    G = np.zeros(F * B, np.float32)
    G[f * B + b] = 1.0
    J, _ = check_full(shape, E, A, G)
    nz = np.flatnonzero(J != 0x88)
    assert list(nz) == [((b * F + f) * 2 + p) * T + t]
    assert bo.dec(int(J[nz[0]])) == (0, 4)  # (-1 + i)(2 - 2i) = 0 + 4i


# ---------------------------------------------------------------- API
def test_host_entry_points():
    T, F, D, B = 64, 2, 64, 16
    shape = bf.Shape.make(T, F, D, B)
    E, A, G = seeded(T, F, D, B, 31)
    want = bo.form_beams(E, A, G, T, F, D, B)
    assert np.array_equal(bf.form_beams(E, A, G, shape), want)
    with bf.Plan(shape, A, G) as plan:
        assert np.array_equal(plan.form_beams_host(E), want)


@pytest.mark.parametrize("e_off,j_off", [(1, 0), (0, 3), (7, 5), (13, 1)])
@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_unaligned_buffers_take_the_mfma_kernel(e_off, j_off, layout):
    # E and J at odd byte addresses: the MFMA kernel (unaligned vector loads
    # and stores), bit-exact, with the bytes around J untouched
    T, F, D, B = 777, 3, 512, 96
    shape = bf.Shape.make(T, F, D, B, j_layout=layout)
    E, A, G = seeded(T, F, D, B, 71 + e_off)
    with bf.Plan(shape, A, G) as plan:
        assert plan.describe().startswith("mfma")
        dE = torch.zeros(shape.e_bytes + 32, dtype=torch.uint8, device="cuda")
        dE[e_off:e_off + shape.e_bytes] = torch.from_numpy(E).cuda()
        dJ = torch.zeros(shape.j_bytes + 32, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        plan.form_beams_device(dJ.data_ptr() + j_off, dE.data_ptr() + e_off, stream())
        torch.cuda.synchronize()
        J = dJ.cpu().numpy()
    assert (J[:j_off] == 0).all() and (J[j_off + shape.j_bytes:] == 0).all(), "bytes outside J were written"
    assert np.array_equal(J[j_off:j_off + shape.j_bytes], bo.form_beams(E, A, G, T, F, D, B, jlayout=layout))


def test_host_path_reused_across_calls():
    # the plan keeps the host path's streams, events and slab buffers: calls
    # after the first reuse them, with new inputs each time
    T, F, D, B = 3000, 2, 256, 64
    shape = bf.Shape.make(T, F, D, B)
    _, A, G = seeded(T, F, D, B, 81)
    with bf.Plan(shape, A, G) as plan:
        for k in range(3):
            E = bo.fill_uniform(T * F * D * 2, 82 + k)
            assert np.array_equal(plan.form_beams_host(E), bo.form_beams(E, A, G, T, F, D, B)), f"call {k}"


@pytest.mark.parametrize("T,F,D,B,layout,pinned", [
    (1000, 3, 64, 40, bf.BF_J_BFPT, False),    # 4 slabs of 256 samples, the last 232
    (4099, 2, 128, 96, bf.BF_J_BFTP, True),    # 8 slabs of 528, the last 403; pinned buffers
    (700, 2, 12, 5, bf.BF_J_BFPT, False),      # generic kernel (D not a multiple of 8)
    (700, 2, 12, 5, bf.BF_J_BFTP, True),
    (255, 1, 32, 16, bf.BF_J_BFPT, False),     # one slab
])
def test_host_path_slabs(T, F, D, B, layout, pinned):
    # the host entry point runs the time axis in slabs on two streams; every
    # slab's J lands in its columns of the caller's J
    say(f"host path T={T} F={F} D={D} B={B} layout={layout} pinned={pinned}")
    shape = bf.Shape.make(T, F, D, B, j_layout=layout)
    E, A, G = seeded(T, F, D, B, 61 + T)
    want = bo.form_beams(E, A, G, T, F, D, B, jlayout=layout)
    with bf.Plan(shape, A, G) as plan:
        if pinned:
            hE = torch.from_numpy(E).pin_memory()
            hJ = torch.empty(shape.j_bytes, dtype=torch.uint8).pin_memory()
            plan.form_beams_host_ptr(hJ.data_ptr(), hE.data_ptr())
            got = hJ.numpy()
        else:
            got = plan.form_beams_host(E)
    assert np.array_equal(got, want)
    assert np.array_equal(bf.form_beams(E, A, G, shape), want)


def test_update_weights_switches_kernels():
    T, F, D, B = 128, 2, 64, 32
    shape = bf.Shape.make(T, F, D, B)
    E, A, G = seeded(T, F, D, B, 41)
    A2 = bo.fill_uniform(F * B * D, 99)
    with bf.Plan(shape, A, G) as plan:
        J, _ = run_device(shape, E, A, G, plan)
        assert np.array_equal(J, bo.form_beams(E, A, G, T, F, D, B))
        plan.update_weights(A2, G)
        J, _ = run_device(shape, E, A2, G, plan)
        assert np.array_equal(J, bo.form_beams(E, A2, G, T, F, D, B))
        Gbig = np.full(F * B, 1e12, np.float32)
        plan.update_weights(A2, Gbig)
        assert plan.describe().startswith("generic")
        J, _ = run_device(shape, E, A2, Gbig, plan)
        assert np.array_equal(J, bo.form_beams(E, A2, Gbig, T, F, D, B))


def test_concurrent_streams_same_plan():
    T, F, D, B = 2048, 4, 512, 96
    shape = bf.Shape.make(T, F, D, B)
    E1, A, G = seeded(T, F, D, B, 51)
    E2 = bo.fill_uniform(E1.size, 52)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with bf.Plan(shape, A, G) as plan:
        dE1, dE2 = torch.from_numpy(E1).cuda(), torch.from_numpy(E2).cuda()
        dJ1 = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
        dJ2 = torch.zeros_like(dJ1)
        torch.cuda.synchronize()
        for _ in range(3):
            plan.form_beams_device(dJ1.data_ptr(), dE1.data_ptr(), s1.cuda_stream)
            plan.form_beams_device(dJ2.data_ptr(), dE2.data_ptr(), s2.cuda_stream)
        torch.cuda.synchronize()
    assert np.array_equal(dJ1.cpu().numpy(), bo.form_beams(E1, A, G, T, F, D, B))
    assert np.array_equal(dJ2.cpu().numpy(), bo.form_beams(E2, A, G, T, F, D, B))


@pytest.mark.parametrize("n,offset", [(0, 0), (1, 0), (15, 3), (16, 0), (1000, 5), (65537, 1), (1 << 22, 7)])
def test_adler32_device(n, offset):
    buf = torch.randint(0, 256, (n + 32,), dtype=torch.uint8, device="cuda")
    got = bf.adler32_device(buf.data_ptr() + offset, n, stream())
    assert got == bo.adler32(buf.cpu().numpy()[offset:offset + n])


@pytest.mark.parametrize("n,offset,ptr_off", [(1, 0, 0), (17, 3, 1), (4096, 100, 0), (100003, 7, 5)])
def test_fill_uniform_device_matches_oracle(n, offset, ptr_off):
    buf = torch.zeros(n + 16, dtype=torch.uint8, device="cuda")
    bf.fill_uniform_device(buf.data_ptr() + ptr_off, n, 1234, offset, stream())
    torch.cuda.synchronize()
    host = buf.cpu().numpy()
    assert np.array_equal(host[ptr_off:ptr_off + n], bo.fill_uniform(n, 1234, offset))
    assert (host[:ptr_off] == 0).all() and (host[ptr_off + n:] == 0).all()
