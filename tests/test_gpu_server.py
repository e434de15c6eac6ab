"""GPU tests of the streaming server (continuous batching, HBM-budget
admission). Every result is compared bit for bit with the oracle run on that
chunk alone, and a chunk's J is shown to be the same alone and in a mix."""
import sys
import time

import numpy as np
import pytest
import torch

import beamforming_amd as bf
from oracle import bf_oracle as bo

pytestmark = pytest.mark.gpu

C, D, B = 20, 512, 96


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def table():
    A = bo.fill_uniform(C * B * D, 31)
    G = bo.setup_gain(C * B, D) * np.float32(8)
    shape = bf.Shape.make(1, C, D, B)
    plan = bf.Plan(shape, A, G)
    yield A, G, plan
    plan.close()


def oracle_request(E, A, G, T, channel, nf):
    Ac = A.reshape(C, B * D)[channel:channel + nf]
    Gc = G.reshape(C, B)[channel:channel + nf]
    return bo.form_beams(E, Ac, Gc, T, nf, D, B)


class Req:
    def __init__(self, T, channel, nf, seed):
        self.T, self.channel, self.nf = T, channel, nf
        self.E = bo.fill_uniform(T * nf * D * 2, seed)
        self.dE = torch.from_numpy(self.E).cuda()
        self.j_bytes = B * nf * 2 * T


def random_reqs(n, seed, tmax=5000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nf = int(rng.choice([1, 4, 16]))
        out.append(Req(int(rng.integers(1, tmax)), int(rng.integers(0, C - nf + 1)), nf, seed * 1000 + i))
    return out


def collect(server, n, timeout=30.0):
    """Waits for n results; returns {tag: (Result, J bytes)} and releases each."""
    got = {}
    while len(got) < n:
        r = server.wait(timeout)
        assert r is not None, f"timed out with {len(got)} of {n} results; stats {server.stats()}"
        assert r.status == 0, f"request {r.tag} failed with status {r.status}"
        got[r.tag] = (r, server.read(r.ticket, r.j_bytes))
        server.release(r.ticket)
    return got


@pytest.mark.parametrize("poll_us", [0, 50])
def test_server_results_match_oracle(table, poll_us):
    # poll_us 0: the loop polls launches in flight without sleeping; 50: timed waits
    A, G, plan = table
    reqs = random_reqs(24, 1)
    torch.cuda.synchronize()
    with bf.Server(plan, hbm_budget=1 << 30, poll_us=poll_us) as srv:
        for i, q in enumerate(reqs):
            srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=i)
        got = collect(srv, len(reqs))
        st = srv.stats()
        say(f"stats {st}")
    assert st["completed"] == len(reqs) and st["released"] == len(reqs) and st["bytes_in_use"] == 0
    for i, q in enumerate(reqs):
        r, J = got[i]
        assert r.t_submit <= r.t_admit <= r.t_launch <= r.t_done
        want = oracle_request(q.E, A, G, q.T, q.channel, q.nf)
        assert np.array_equal(J, want), f"request {i} (T={q.T}, channel={q.channel}, nf={q.nf})"


@pytest.mark.parametrize("use_graph", [False, True])
def test_alone_equals_in_mix(table, use_graph):
    A, G, plan = table
    probe = Req(3001, 2, 4, 777)
    others = random_reqs(40, 2)
    torch.cuda.synchronize()
    with bf.Server(plan, hbm_budget=2 << 30, use_graph=use_graph) as srv:
        srv.submit(probe.dE.data_ptr(), probe.T, probe.channel, probe.nf, tag=1000)
        alone = collect(srv, 1)[1000]
        assert alone[0].batch_size == 1
        # the mix: one atomic burst, so that the launch carries many requests
        # at once, with the probe in the middle of it; then single requests
        # arriving while that launch runs
        burst = [(q.dE.data_ptr(), q.T, q.channel, q.nf, i) for i, q in enumerate(others[:30])]
        burst.insert(15, (probe.dE.data_ptr(), probe.T, probe.channel, probe.nf, 1001))
        srv.submit_many(burst)
        for i, q in enumerate(others[30:]):
            srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=30 + i)
        mix = collect(srv, len(others) + 1)
        st = srv.stats()
        say(f"graph={use_graph} stats {st} probe batch {mix[1001][0].batch_size}")
    assert mix[1001][0].batch_size > 1, "the probe was not batched with other requests"
    assert st["max_batch"] > 1
    assert np.array_equal(alone[1], mix[1001][1])
    assert np.array_equal(alone[1], oracle_request(probe.E, A, G, probe.T, probe.channel, probe.nf))
    for i, q in enumerate(others):
        assert np.array_equal(mix[i][1], oracle_request(q.E, A, G, q.T, q.channel, q.nf)), f"request {i}"


def test_budget_admission_and_release(table):
    A, G, plan = table
    reqs = [Req(4096, c, 1, 50 + c) for c in range(10)]
    jb = reqs[0].j_bytes  # 786432 bytes each
    budget = 2 * jb + 1000  # room for two results at a time
    torch.cuda.synchronize()
    with bf.Server(plan, hbm_budget=budget) as srv:
        for i, q in enumerate(reqs):
            srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=i)
        time.sleep(0.2)
        st = srv.stats()
        say(f"before release: {st}")
        # no result has been released, so only two can have been admitted
        assert st["admitted"] == 2 and st["waiting"] == 8
        got = collect(srv, len(reqs))
        st = srv.stats()
        say(f"after: {st}")
        assert st["peak_bytes"] <= budget and st["bytes_in_use"] == 0
        with pytest.raises(bf.BeamformError) as e:  # alone larger than the whole budget
            srv.submit(reqs[0].dE.data_ptr(), 16384, 0, 1)
        assert e.value.status == bf.BF_ERR_OVER_BUDGET
    for i, q in enumerate(reqs):
        assert np.array_equal(got[i][1], oracle_request(q.E, A, G, q.T, q.channel, q.nf))


def test_host_fed_requests(table):
    A, G, plan = table
    reqs = random_reqs(6, 3, tmax=3000)
    with bf.Server(plan, hbm_budget=1 << 28) as srv:
        for i, q in enumerate(reqs):
            srv.submit(q.E.ctypes.data, q.T, q.channel, q.nf, tag=i, e_on_host=True)
        got = collect(srv, len(reqs))
        assert srv.stats()["bytes_in_use"] == 0
    for i, q in enumerate(reqs):
        assert np.array_equal(got[i][1], oracle_request(q.E, A, G, q.T, q.channel, q.nf))


def test_bad_requests_and_shutdown(table):
    A, G, plan = table
    q = Req(64, 0, 1, 9)
    torch.cuda.synchronize()
    with pytest.raises(bf.BeamformError):
        bf.Server(plan, hbm_budget=0)
    for poll_us in (-1, 1000001):
        with pytest.raises(bf.BeamformError):
            bf.Server(plan, hbm_budget=1 << 20, poll_us=poll_us)
    srv = bf.Server(plan, hbm_budget=1 << 26)
    for bad in [dict(ntimes=0, channel=0, nfreq=1), dict(ntimes=64, channel=C, nfreq=1),
                dict(ntimes=64, channel=C - 2, nfreq=4), dict(ntimes=64, channel=-1, nfreq=1)]:
        with pytest.raises(bf.BeamformError) as e:
            srv.submit(q.dE.data_ptr(), bad["ntimes"], bad["channel"], bad["nfreq"])
        assert e.value.status == bf.BF_ERR_INVALID_ARGUMENT
    # shut down with work queued and results unreleased: must return promptly
    big = Req(16384, 0, 4, 10)
    torch.cuda.synchronize()
    for i in range(8):
        srv.submit(big.dE.data_ptr(), big.T, big.channel, big.nf, tag=i)
    t0 = time.time()
    srv.close()
    assert time.time() - t0 < 10


def test_weight_updates_while_serving(table):
    # Requests carry the weights current when they were submitted: three
    # weight sets, each update issued right after a burst, with the earlier
    # bursts still queued or running.
    A0, G0, plan = table
    sets = [(A0, G0)]
    for k in (1, 2):
        sets.append((bo.fill_uniform(C * B * D, 40 + k), bo.setup_gain(C * B, D) * np.float32(4 + 2 * k)))
    bursts = [random_reqs(10, 60 + k, tmax=8000) for k in range(3)]
    torch.cuda.synchronize()
    with bf.Server(plan, hbm_budget=1 << 30) as srv:
        for k, reqs in enumerate(bursts):
            if k > 0:
                srv.update_weights(*sets[k])
            for i, q in enumerate(reqs):
                srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=100 * k + i)
        got = collect(srv, sum(len(r) for r in bursts))
        srv.update_weights(*sets[0])  # retires the finished sets
        q0 = bursts[0][0]
        srv.submit(q0.dE.data_ptr(), q0.T, q0.channel, q0.nf, tag=999)
        again = collect(srv, 1)
    for k, reqs in enumerate(bursts):
        A, G = sets[k]
        for i, q in enumerate(reqs):
            want = oracle_request(q.E, A, G, q.T, q.channel, q.nf)
            assert np.array_equal(got[100 * k + i][1], want), f"set {k} request {i}"
    A, G = sets[0]
    assert np.array_equal(again[999][1], oracle_request(q0.E, A, G, q0.T, q0.channel, q0.nf))


def test_weight_update_refused_outside_mfma_path(table):
    A, G, plan = table
    with bf.Server(plan, hbm_budget=1 << 28) as srv:
        bad = G.copy()
        bad[0] = np.float32(np.inf)  # the plan would fall back to the generic kernel
        with pytest.raises(bf.BeamformError):
            srv.update_weights(A, bad)
        q = random_reqs(1, 77)[0]
        srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=1)
        got = collect(srv, 1)
    assert np.array_equal(got[1][1], oracle_request(q.E, A, G, q.T, q.channel, q.nf))


def test_submitters_on_several_threads(table):
    # four host threads submit at once (the library takes its lock; ctypes
    # releases the GIL during each call) while the main thread collects
    import threading
    A, G, plan = table
    groups = [random_reqs(8, 90 + k, tmax=3000) for k in range(4)]
    torch.cuda.synchronize()
    errors = []
    with bf.Server(plan, hbm_budget=1 << 29, max_inflight=16) as srv:
        def submit_all(k):
            try:
                for i, q in enumerate(groups[k]):
                    srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=100 * k + i)
            except Exception as exc:  # reported below
                errors.append(exc)
        threads = [threading.Thread(target=submit_all, args=(k,)) for k in range(4)]
        for t in threads:
            t.start()
        got = collect(srv, 32)
        for t in threads:
            t.join()
        st = srv.stats()
    assert not errors, errors
    assert st["submitted"] == 32 and st["completed"] == 32 and st["bytes_in_use"] == 0
    for k, reqs in enumerate(groups):
        for i, q in enumerate(reqs):
            want = oracle_request(q.E, A, G, q.T, q.channel, q.nf)
            assert np.array_equal(got[100 * k + i][1], want), f"thread {k} request {i}"


def test_submitters_race_the_completion_scan(table, monkeypatch):
    # A submitting thread (poll_us 0) takes a launch slot, drops the server
    # lock and only then enqueues the launch and records its completion
    # event; the loop thread's completion scan must leave such a slot alone
    # (before the event is recorded it reads as complete). The library's
    # BF_SERVER_ISSUE_DELAY_US pause widens that window to 3 ms, and four
    # threads submit at once: a launch finished early would hand out J before
    # the kernel wrote it, which the oracle comparison catches.
    import threading
    monkeypatch.setenv("BF_SERVER_ISSUE_DELAY_US", "3000")
    A, G, plan = table
    groups = [random_reqs(6, 120 + k, tmax=3000) for k in range(4)]
    torch.cuda.synchronize()
    errors = []
    with bf.Server(plan, hbm_budget=1 << 29, poll_us=0) as srv:
        def submit_all(k):
            try:
                for i, q in enumerate(groups[k]):
                    srv.submit(q.dE.data_ptr(), q.T, q.channel, q.nf, tag=100 * k + i)
            except Exception as exc:  # reported below
                errors.append(exc)
        threads = [threading.Thread(target=submit_all, args=(k,)) for k in range(4)]
        for t in threads:
            t.start()
        got = collect(srv, 24)
        for t in threads:
            t.join()
        st = srv.stats()
    say(f"stats {st}")
    assert not errors, errors
    assert st["completed"] == 24 and st["failed"] == 0
    for k, reqs in enumerate(groups):
        for i, q in enumerate(reqs):
            want = oracle_request(q.E, A, G, q.T, q.channel, q.nf)
            assert np.array_equal(got[100 * k + i][1], want), f"thread {k} request {i}"


def test_unaligned_requests(table):
    A, G, plan = table
    q = Req(2500, 3, 4, 88)
    dE = torch.zeros(q.E.size + 32, dtype=torch.uint8, device="cuda")
    dE[5:5 + q.E.size] = q.dE
    torch.cuda.synchronize()
    with bf.Server(plan, hbm_budget=1 << 28) as srv:
        srv.submit(dE.data_ptr() + 5, q.T, q.channel, q.nf, tag=1)
        got = collect(srv, 1)
    assert np.array_equal(got[1][1], oracle_request(q.E, A, G, q.T, q.channel, q.nf))


def test_j2_layout_plan_through_the_server():
    # a channel table with the [b][f][t][p] output layout: lone and batched
    # launches both write J2, bit-exact with the oracle
    A = bo.fill_uniform(C * B * D, 51)
    G = bo.setup_gain(C * B, D) * np.float32(8)
    reqs = random_reqs(12, 7)
    torch.cuda.synchronize()
    with bf.Plan(bf.Shape.make(1, C, D, B, j_layout=bf.BF_J_BFTP), A, G) as plan:
        with bf.Server(plan, hbm_budget=1 << 30) as srv:
            srv.submit(reqs[0].dE.data_ptr(), reqs[0].T, reqs[0].channel, reqs[0].nf, tag=0)  # alone
            first = collect(srv, 1)
            srv.submit_many([(q.dE.data_ptr(), q.T, q.channel, q.nf, i) for i, q in enumerate(reqs) if i > 0])
            got = collect(srv, len(reqs) - 1)
            got.update(first)
    for i, q in enumerate(reqs):
        Ac = A.reshape(C, B * D)[q.channel:q.channel + q.nf]
        Gc = G.reshape(C, B)[q.channel:q.channel + q.nf]
        want = bo.form_beams(q.E, Ac, Gc, q.T, q.nf, D, B, jlayout=1)
        assert np.array_equal(got[i][1], want), f"request {i} (T={q.T}, channel={q.channel}, nf={q.nf})"
