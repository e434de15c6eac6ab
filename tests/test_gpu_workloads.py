"""GPU parity at the batched and serving workloads themselves (not scaled-down
stand-ins), every output compared with the oracle (the CPU restatement of the
reference cpu.cxx:15-46, run on each chunk or request alone):

* batched: 64 channels x D=512 x B=96 x 2 polarizations, each channel with
  its own ntimes drawn log-uniform in [512, 32768] (seed 7, sum 470 964),
  one persistent launch per run, eager and hipGraph replay, with the
  autotuned work-item size and both item modes (balanced items, one per
  workgroup, and fixed-size items from the device-side queue);
* serving: 200 requests arriving as a seeded Poisson process at 2000 chunks/s,
  nfreq in {1, 4, 16}, ntimes uniform in [1024, 16384], D=512, B=96, weights
  resident, HBM-budget admission; then a probe request alone and the same
  probe inside a burst with all 200 requests (alone == in the mix, exactly).

The workloads come from bench.py's own generators, so the benchmark times
exactly the inputs checked here."""
import hashlib
import os
import sys
import threading
import time

import numpy as np
import pytest
import torch

import beamforming_amd as bf
from oracle import bf_oracle as bo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

D, B = 512, 96


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stream():
    return torch.cuda.current_stream().cuda_stream


def digest(a):
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=16).hexdigest()


def oracle_chunk(E, A, G, C, channel, nf, T):
    Ac = A.reshape(C, B * D)[channel:channel + nf]
    Gc = G.reshape(C, B)[channel:channel + nf]
    return bo.form_beams(E, Ac, Gc, T, nf, D, B)


# ---------------------------------------------------------------------------
# batched workload

@pytest.fixture(scope="module")
def batched():
    C, seed = 64, 7
    Ts = bench.batched_chunk_lengths(C, seed)
    assert sum(Ts) == 470964 and min(Ts) >= 512 and max(Ts) <= 32768
    A = bo.fill_uniform(C * B * D, seed + 1)
    G = bo.setup_gain(C * B, D)
    e_off, e_total = bench.aligned_offsets([T * D * 2 for T in Ts])
    j_off, j_total = bench.aligned_offsets([B * 2 * T for T in Ts])
    dE = torch.empty(e_total, dtype=torch.uint8, device="cuda")
    bf.fill_uniform_device(dE.data_ptr(), e_total, seed + 2, 0, stream())
    dJ = torch.empty(j_total, dtype=torch.uint8, device="cuda")
    E = dE.cpu().numpy()
    say(f"batched workload: {C} chunks, sum ntimes {sum(Ts)}; oracle ...")
    t0 = time.time()
    want = [oracle_chunk(E[e_off[c]:e_off[c] + Ts[c] * D * 2], A, G, C, c, 1, Ts[c]) for c in range(C)]
    say(f"  oracle {time.time() - t0:.1f} s")
    yield dict(C=C, Ts=Ts, A=A, G=G, dE=dE, dJ=dJ, e_off=e_off, j_off=j_off, want=want)


def check_batched(w, what):
    got = w["dJ"].cpu().numpy()
    for c, T in enumerate(w["Ts"]):
        J = got[w["j_off"][c]:w["j_off"][c] + B * 2 * T]
        bad = np.flatnonzero(J != w["want"][c])
        assert bad.size == 0, f"{what}: chunk {c} (ntimes {T}): {bad.size} bytes differ from the oracle"


@pytest.mark.parametrize("mode", ["autotuned", "balanced", "queue"])
def test_batched_workload_matches_oracle(batched, mode):
    w = batched
    C = w["C"]
    chunks = [(w["dE"].data_ptr() + w["e_off"][c], w["dJ"].data_ptr() + w["j_off"][c], w["Ts"][c], c)
              for c in range(C)]
    with bf.Plan(bf.Shape.make(1, C, D, B), w["A"], w["G"]) as plan, bf.Batch(plan, C) as batch:
        if mode == "autotuned":
            say(batch.autotune(chunks, stream(), iters=2).strip().splitlines()[-1])
        else:
            batch.set_item_tiles(0 if mode == "balanced" else 64)
        for graph in (False, True):
            for rep in range(2 if graph else 1):
                w["dJ"].zero_()  # 0x00 decodes to (-8, -8), which form_beams never writes
                torch.cuda.synchronize()
                batch.run(chunks, stream(), graph=graph)
                torch.cuda.synchronize()
                say(f"  {mode} graph={graph} rep={rep}: {batch.describe()}")
                check_batched(w, f"{mode} graph={graph} replay {rep}")
                if graph:  # eager work between replays
                    batch.run(chunks[:3], stream())
                    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# serving workload

def test_serving_workload_matches_oracle():
    C, n, rate, seed = 64, 200, 2000.0, 9
    pool_bytes = 512 << 20  # holds the largest request's E (16 channels x 16384 samples = 256 MiB)
    rng = np.random.default_rng(seed)
    arrivals, Ts, nfs, chs, offs = bench.serving_requests(n, rate, rng, C, D, pool_bytes)
    assert Ts.min() >= 1024 and Ts.max() <= 16384 and set(nfs.tolist()) <= {1, 4, 16}
    A = bo.fill_uniform(C * B * D, seed + 1)
    G = bo.setup_gain(C * B, D)
    pool = torch.empty(pool_bytes, dtype=torch.uint8, device="cuda")
    bf.fill_uniform_device(pool.data_ptr(), pool_bytes, seed + 2, 0, stream())
    pool_host = pool.cpu().numpy()
    reqs = [(pool.data_ptr() + int(offs[i]), int(Ts[i]), int(chs[i]), int(nfs[i]), i) for i in range(n)]
    probe = 1000  # tag of the probe: request 0's chunk again

    with bf.Plan(bf.Shape.make(1, C, D, B), A, G) as plan, bf.Server(plan, hbm_budget=8 << 30) as srv:
        got, meta = {}, {}

        def consume(count, phase):
            while sum(1 for k in got if k[0] == phase) < count:
                r = srv.wait(30.0)
                assert r is not None, f"server stalled: {srv.stats()}"
                assert r.status == 0, f"request {r.tag} failed: status {r.status}"
                got[(phase, r.tag)] = digest(srv.read(r.ticket, r.j_bytes))
                meta[(phase, r.tag)] = r.batch_size
                srv.release(r.ticket)

        # phase 1: Poisson arrivals at the configured rate
        th = threading.Thread(target=consume, args=(n, 1))
        th.start()
        t0 = time.perf_counter()
        for i, q in enumerate(reqs):
            due = t0 + arrivals[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                if due - now > 2e-4:
                    time.sleep(5e-5)  # lets the consumer thread take the GIL
            srv.submit(q[0], q[1], q[2], q[3], tag=q[4])
        th.join()
        st1 = srv.stats()
        say(f"poisson phase: {st1}")
        # phase 2: the probe alone
        srv.submit(reqs[0][0], reqs[0][1], reqs[0][2], reqs[0][3], tag=probe)
        consume(1, 2)
        assert meta[(2, probe)] == 1
        # phase 3: the probe in the middle of one burst of all requests
        burst = list(reqs)
        burst.insert(n // 2, (reqs[0][0], reqs[0][1], reqs[0][2], reqs[0][3], probe))
        srv.submit_many(burst)
        consume(n + 1, 3)
        say(f"burst phase: {srv.stats()}; probe launched with {meta[(3, probe)]} requests")
    assert meta[(3, probe)] > 1, "the probe was not batched with other requests"
    assert got[(3, probe)] == got[(2, probe)], "the probe's J differs alone and in the mix"

    t0 = time.time()
    for i in range(n):
        T, c, nf = int(Ts[i]), int(chs[i]), int(nfs[i])
        E = pool_host[int(offs[i]):int(offs[i]) + T * nf * D * 2]
        want = digest(oracle_chunk(E, A, G, C, c, nf, T))
        assert got[(1, i)] == want, f"poisson phase: request {i} (ntimes {T}, channel {c}, nfreq {nf}) differs"
        assert got[(3, i)] == want, f"burst phase: request {i} (ntimes {T}, channel {c}, nfreq {nf}) differs"
        if i == 0:
            assert got[(2, probe)] == want
    say(f"oracle over {n} requests: {time.time() - t0:.1f} s")
