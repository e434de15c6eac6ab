"""GPU tests of E ingest: packets E[t][F][D][P] (reference Elinear) become
per-channel-group chunks [T_chunk][G][D][P] in device ring buffers. The
chunks are checked byte for byte against the packets' slices, and beams
formed from them through the batch path against the oracle."""
import numpy as np
import pytest
import torch

import beamforming_amd as bf
from oracle import bf_oracle as bo

pytestmark = pytest.mark.gpu


def stream():
    return torch.cuda.current_stream().cuda_stream


def device_bytes(ptr, n):
    """Copies n bytes of device memory at ptr to a new host array (HIP runtime)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    host = np.empty(n, dtype=np.uint8)
    assert hip.hipMemcpy(host.ctypes.data, ctypes.c_void_p(ptr), n, 2) == 0  # hipMemcpyDeviceToHost
    return host


@pytest.mark.parametrize("mode", ["host", "device", "device_unaligned"])
def test_ingest_chunks_are_packet_slices(mode):
    # device packets go through the scatter kernel; host packets and device
    # packets that are not 16-byte aligned through 2-D copies
    on_host = mode == "host"
    F, D, G, TC = 8, 64, 4, 100
    total = 3 * TC + 57
    E = bo.fill_uniform(total * F * D * 2, 71).reshape(total, F, D * 2)
    sizes = [1, 17, 100, 64, 99, 3, 53]  # ragged packets crossing chunk boundaries
    sizes.append(total - sum(sizes))
    assert sizes[-1] > 0
    shift = 8 if mode == "device_unaligned" else 0
    dE_buf = torch.zeros(E.size + 16, dtype=torch.uint8, device="cuda")
    dE_buf[shift:shift + E.size].copy_(torch.from_numpy(E.reshape(-1)))
    base = dE_buf.data_ptr() + shift
    got = []
    with bf.Ingest(F, D, TC, G, ring_slots=4) as ing:
        t = 0
        for n in sizes:
            if on_host:
                pkt = np.ascontiguousarray(E[t:t + n]).reshape(-1)
                assert ing.push(pkt.ctypes.data, n, True)
                ing.sync()  # the packet buffer may be reused from here on
            else:
                assert ing.push(base + t * F * D * 2, n, False)
            t += n
            while True:
                r = ing.pop(stream())
                if r is None:
                    break
                c, sid = r
                torch.cuda.synchronize()
                got.append((c.channel, c.nfreq, c.ntimes, device_bytes(c.E, TC * G * D * 2)))
                ing.release(sid, stream())
        filled, ready = ing.pending()
        assert filled == total % TC and ready == 0
    assert len(got) == (total // TC) * (F // G)
    for i, (ch, nf, nt, data) in enumerate(got):
        k = i // (F // G)  # chunk index in time
        assert nf == G and nt == TC and ch == (i % (F // G)) * G
        want = E[k * TC:(k + 1) * TC, ch:ch + G].reshape(-1)
        assert np.array_equal(data, want), f"chunk {i}"


def test_ingest_backpressure_and_beams():
    F, D, B, G, TC = 8, 512, 96, 4, 256
    A = bo.fill_uniform(F * B * D, 72)
    Gn = bo.setup_gain(F * B, D) * np.float32(8)
    E = bo.fill_uniform(4 * TC * F * D * 2, 73).reshape(4 * TC, F, D * 2)
    dE = torch.from_numpy(E.reshape(-1)).cuda()
    with bf.Plan(bf.Shape.make(1, F, D, B), A, Gn) as plan, bf.Batch(plan, 8) as batch, \
            bf.Ingest(F, D, TC, G, ring_slots=2) as ing:
        # two chunks fill both slots of every group; the third needs a released slot
        assert ing.push(dE.data_ptr(), 2 * TC, False)
        popped = [ing.pop(stream()) for _ in range(2 * (F // G))]
        assert all(p is not None for p in popped) and ing.pop(stream()) is None
        assert not ing.push(dE.data_ptr() + 2 * TC * F * D * 2, TC, False)  # no free slot: refused
        assert ing.pending() == (0, 0)
        # beams from the popped chunks, then release their slots after that work
        Js = [torch.zeros(B * G * 2 * TC, dtype=torch.uint8, device="cuda") for _ in popped]
        chunks = [(c.E, J.data_ptr(), c.ntimes, c.channel, c.nfreq) for (c, _), J in zip(popped, Js)]
        batch.run(chunks, stream())
        for _, sid in popped:
            ing.release(sid, stream())
        torch.cuda.synchronize()
        assert ing.push(dE.data_ptr() + 2 * TC * F * D * 2, TC, False)  # slots are free again
        torch.cuda.synchronize()
    for i, ((c, _), J) in enumerate(zip(popped, Js)):
        k = i // (F // G)
        ch = c.channel
        Ek = np.ascontiguousarray(E[k * TC:(k + 1) * TC, ch:ch + G]).reshape(-1)
        want = bo.form_beams(Ek, A.reshape(F, -1)[ch:ch + G], Gn.reshape(F, -1)[ch:ch + G], TC, G, D, B)
        assert np.array_equal(J.cpu().numpy(), want), f"chunk {i}"


def test_ingest_many_groups():
    # 300 one-channel groups: the scatter kernel runs in launches of up to 256 groups
    F, D, G, TC = 300, 8, 1, 5
    total = 3 * TC + 2
    E = bo.fill_uniform(total * F * D * 2, 74).reshape(total, F, D * 2)
    dE = torch.from_numpy(E.reshape(-1)).cuda()
    got = []
    with bf.Ingest(F, D, TC, G, ring_slots=4) as ing:
        t = 0
        for n in [4, 7, 6]:  # the 7-sample packet fills the open chunk, one more and starts a third
            assert ing.push(dE.data_ptr() + t * F * D * 2, n, False)
            t += n
            while True:
                r = ing.pop(stream())
                if r is None:
                    break
                c, sid = r
                torch.cuda.synchronize()
                got.append((c.channel, c.ntimes, device_bytes(c.E, TC * G * D * 2)))
                ing.release(sid, stream())
        torch.cuda.synchronize()
    assert len(got) == (total // TC) * F
    for i, (ch, nt, data) in enumerate(got):
        k = i // F
        assert nt == TC and ch == i % F
        assert np.array_equal(data, E[k * TC:(k + 1) * TC, ch].reshape(-1)), f"chunk {i}"


def test_push_many_and_partial_refusal():
    # packets pushed several per call: the chunks are the packets' slices; with
    # two ring slots and nothing released, the call stops at the first packet
    # that needs a third slot and reports how many it took
    F, D, G, TC = 8, 64, 2, 64
    total = 3 * TC
    E = bo.fill_uniform(total * F * D * 2, 74).reshape(total, F, D * 2)
    dE = torch.from_numpy(E.reshape(-1)).cuda()
    row = F * D * 2
    starts = list(range(0, total, 16))
    with bf.Ingest(F, D, TC, G, ring_slots=2) as ing:
        k = ing.push_many([dE.data_ptr() + t * row for t in starts], [16] * len(starts), False)
        assert k == 2 * TC // 16  # the third chunk has no free slot
        ing.sync()
        got = []
        while True:
            r = ing.pop(stream())
            if r is None:
                break
            c, sid = r
            torch.cuda.synchronize()
            got.append((c.channel, device_bytes(c.E, TC * G * D * 2)))
        assert len(got) == 2 * (F // G)
        for i, (ch, data) in enumerate(got):
            kk = i // (F // G)
            assert np.array_equal(data, E[kk * TC:(kk + 1) * TC, ch:ch + G].reshape(-1)), f"chunk {i}"


def test_wait_release_frees_the_oldest_slot():
    # two ring slots, both chunk times filled and popped: a push into a third
    # chunk time is refused until a release completes; wait_release blocks on
    # the oldest release only, after which the push goes through. With no
    # release pending it returns False at once.
    F, D, G, TC = 8, 64, 2, 64
    total = 4 * TC
    E = bo.fill_uniform(total * F * D * 2, 76).reshape(total, F, D * 2)
    dE = torch.from_numpy(E.reshape(-1)).cuda()
    row = F * D * 2
    ng = F // G
    with bf.Ingest(F, D, TC, G, ring_slots=2) as ing:
        assert not ing.wait_release()  # nothing released yet
        assert ing.push_many([dE.data_ptr(), dE.data_ptr() + TC * row], [TC, TC], False) == 2
        assert not ing.push(dE.data_ptr() + 2 * TC * row, TC, False)  # both slots held
        popped = []
        while True:
            r = ing.pop(stream())
            if r is None:
                break
            popped.append(r)
        assert len(popped) == 2 * ng
        for c, sid in popped:
            assert c.ntimes == TC
        # release the first chunk time's slots, the second's stay held
        for c, sid in popped[:ng]:
            ing.release(sid, stream())
        waited = 0
        while not ing.push(dE.data_ptr() + 2 * TC * row, TC, False):
            assert ing.wait_release(), "a release is pending"
            waited += 1
            assert waited <= ng
        # the second chunk time's slots are still held: the fourth chunk time
        # is refused and no release is pending
        assert not ing.wait_release()
        assert not ing.push(dE.data_ptr() + 3 * TC * row, TC, False)
        for c, sid in popped[ng:]:
            ing.release(sid, stream())
        while not ing.push(dE.data_ptr() + 3 * TC * row, TC, False):
            assert ing.wait_release()
        ing.sync()
        got = []
        while True:
            r = ing.pop(stream())
            if r is None:
                break
            torch.cuda.synchronize()
            got.append((r[0].channel, device_bytes(r[0].E, TC * G * D * 2)))
        assert len(got) == 2 * ng
        for i, (ch, data) in enumerate(got):
            kk = 2 + i // ng
            assert np.array_equal(data, E[kk * TC:(kk + 1) * TC, ch:ch + G].reshape(-1)), f"chunk {i}"
