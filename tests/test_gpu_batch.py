"""GPU parity of the channel-table paths: one-channel launches, the batched
persistent launch (plain and hipGraph replay) and the autotuners. Every chunk
is compared bit for bit with the oracle run on that chunk alone (F = 1 with
the channel's weights and gains)."""
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


def channel_table(C, D, B, seed, gain_scale=8.0):
    A = bo.fill_uniform(C * B * D, seed)
    G = bo.setup_gain(C * B, D) * np.float32(gain_scale)
    return A, G


def oracle_chunk(E, A, G, channel, T, D, B, layout=0):
    Ac = A.reshape(-1, B * D)[channel]
    Gc = G.reshape(-1, B)[channel]
    return bo.form_beams(E, Ac, Gc, T, 1, D, B, jlayout=layout)


def make_chunks(specs, D, B, seed):
    """specs: list of (ntimes, channel). Returns host E arrays, device E/J tensors."""
    host, dev = [], []
    for i, (T, c) in enumerate(specs):
        E = bo.fill_uniform(T * D * 2, seed + 17 * i)
        dE = torch.from_numpy(E).cuda()
        dJ = torch.zeros(B * 2 * T, dtype=torch.uint8, device="cuda")
        host.append(E)
        dev.append((dE, dJ))
    return host, dev


def chunk_list(specs, dev):
    return [(dE.data_ptr(), dJ.data_ptr(), T, c) for (T, c), (dE, dJ) in zip(specs, dev)]


def check_chunks(specs, host, dev, A, G, D, B, layout=0):
    for i, ((T, c), E, (dE, dJ)) in enumerate(zip(specs, host, dev)):
        got = dJ.cpu().numpy()
        assert (got != 0).all(), f"chunk {i}: unwritten outputs"
        want = oracle_chunk(E, A, G, c, T, D, B, layout)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"chunk {i} (T={T}, channel={c}): {bad.size} mismatches"


@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_channel_launch(layout):
    C, D, B = 5, 512, 96
    A, G = channel_table(C, D, B, 3)
    shape = bf.Shape.make(1, C, D, B, j_layout=layout)
    specs = [(1000, 0), (16, 4), (3, 2), (4096, 3), (777, 1)]
    host, dev = make_chunks(specs, D, B, 100)
    with bf.Plan(shape, A, G) as plan:
        say(f"channel plan: {plan.describe()}")
        for (T, c), (dE, dJ) in zip(specs, dev):
            plan.form_beams_channel(c, T, dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
    check_chunks(specs, host, dev, A, G, D, B, layout)


@pytest.mark.parametrize("D,B", [(512, 96), (64, 16), (512, 512), (40, 50)])
def test_batch_matches_oracle(D, B):
    C = 7
    A, G = channel_table(C, D, B, 5)
    shape = bf.Shape.make(1, C, D, B)
    rng = np.random.default_rng(D + B)
    specs = [(int(rng.integers(1, 3000)), int(rng.integers(0, C))) for _ in range(11)] + [(1, 0), (16, 6), (4097, 3)]
    host, dev = make_chunks(specs, D, B, 200)
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 32) as batch:
        say(f"batch D={D} B={B}")
        batch.run(chunk_list(specs, dev), stream())
        torch.cuda.synchronize()
        say(f"  {batch.describe()}")
    check_chunks(specs, host, dev, A, G, D, B)


def test_batch_graph_replays_match_eager():
    C, D, B = 8, 512, 96
    A, G = channel_table(C, D, B, 9)
    shape = bf.Shape.make(1, C, D, B)
    rng = np.random.default_rng(4)
    lists = []
    for r in range(3):  # different chunk lists through the same captured graph
        specs = [(int(rng.integers(100, 5000)), int(rng.integers(0, C))) for _ in range(6 + 3 * r)]
        lists.append(specs)
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 16) as batch:
        for r, specs in enumerate(lists):
            host, dev = make_chunks(specs, D, B, 300 + 50 * r)
            for rep in range(3):
                for _, dJ in dev:
                    dJ.zero_()
                torch.cuda.synchronize()
                batch.run(chunk_list(specs, dev), stream(), graph=True)
                torch.cuda.synchronize()
                say(f"  replay {r}.{rep}: {batch.describe()}")
                check_chunks(specs, host, dev, A, G, D, B)
                # eager work between replays
                batch.run(chunk_list(specs[:2], dev[:2]), stream())
                torch.cuda.synchronize()


def test_batch_runs_on_two_streams_back_to_back():
    # Two runs of one batch on different streams with no synchronisation in
    # between: they use the batch's two descriptor blocks and may overlap.
    C, D, B = 6, 512, 96
    A, G = channel_table(C, D, B, 13)
    shape = bf.Shape.make(1, C, D, B)
    specs1 = [(20000, 1), (333, 4), (7000, 0)]
    specs2 = [(50, 2), (9000, 5), (1234, 3), (16, 1)]
    host1, dev1 = make_chunks(specs1, D, B, 400)
    host2, dev2 = make_chunks(specs2, D, B, 500)
    s1 = torch.cuda.Stream()
    s2 = torch.cuda.Stream()
    torch.cuda.synchronize()
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 8) as batch:
        say("batch on two streams")
        batch.run(chunk_list(specs1, dev1), s1.cuda_stream)
        batch.run(chunk_list(specs2, dev2), s2.cuda_stream)
        torch.cuda.synchronize()
    check_chunks(specs1, host1, dev1, A, G, D, B)
    check_chunks(specs2, host2, dev2, A, G, D, B)


def test_batch_item_sizes_and_autotune():
    C, D, B = 4, 512, 96
    A, G = channel_table(C, D, B, 11)
    shape = bf.Shape.make(1, C, D, B)
    specs = [(2048, 0), (512, 1), (8192, 2), (100, 3), (3000, 1)]
    host, dev = make_chunks(specs, D, B, 400)
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 8) as batch:
        for tiles in (8, 24, 64, 256, 0):
            batch.set_item_tiles(tiles)
            for _, dJ in dev:
                dJ.zero_()
            batch.run(chunk_list(specs, dev), stream())
            torch.cuda.synchronize()
            check_chunks(specs, host, dev, A, G, D, B)
        report = batch.autotune(chunk_list(specs, dev), stream(), iters=2)
        say(report)
        assert "chosen" in report
        for _, dJ in dev:
            dJ.zero_()
        batch.run(chunk_list(specs, dev), stream())
        torch.cuda.synchronize()
    check_chunks(specs, host, dev, A, G, D, B)


def test_batch_rejects_bad_chunks():
    C, D, B = 2, 64, 16
    A, G = channel_table(C, D, B, 13)
    shape = bf.Shape.make(1, C, D, B)
    dE = torch.zeros(64 * D * 2 + 16, dtype=torch.uint8, device="cuda")
    dJ = torch.zeros(B * 2 * 64 + 16, dtype=torch.uint8, device="cuda")
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 2) as batch:
        with pytest.raises(bf.BeamformError):
            batch.run([(dE.data_ptr(), dJ.data_ptr(), 64, 5)], stream())  # channel out of range
        with pytest.raises(bf.BeamformError):
            batch.run([(dE.data_ptr(), dJ.data_ptr(), 64, 0)] * 3, stream())  # more than max_chunks
        batch.run([], stream())  # empty list is a no-op
        torch.cuda.synchronize()


def test_plan_autotune_keeps_results_exact():
    T, F, D, B = 4096, 4, 512, 96
    shape = bf.Shape.make(T, F, D, B)
    E = bo.fill_uniform(T * F * D * 2, 21)
    A = bo.fill_uniform(F * B * D, 22)
    G = bo.setup_gain(F * B, D) * np.float32(8)
    dE = torch.from_numpy(E).cuda()
    dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
    with bf.Plan(shape, A, G) as plan:
        report = plan.autotune(dJ.data_ptr(), dE.data_ptr(), stream(), iters=2)
        say(report)
        assert "chosen" in report
        dJ.zero_()
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
        say(plan.describe())
    assert np.array_equal(dJ.cpu().numpy(), bo.form_beams(E, A, G, T, F, D, B))


@pytest.mark.parametrize("nt,stage", [(3, True), (3, False), (1, True), (1, False), (2, True), (2, False)])
def test_configured_layouts_are_exact(nt, stage):
    # every beam tiling and LDS layout the autotuner can pick, plain and batched
    T, F, D, B = 1000, 3, 512, 96
    shape = bf.Shape.make(T, F, D, B)
    E = bo.fill_uniform(T * F * D * 2, 23)
    A = bo.fill_uniform(F * B * D, 24)
    G = bo.setup_gain(F * B, D) * np.float32(8)
    want = bo.form_beams(E, A, G, T, F, D, B)
    dE = torch.from_numpy(E).cuda()
    dJ = torch.zeros(shape.j_bytes, dtype=torch.uint8, device="cuda")
    with bf.Plan(shape, A, G) as plan:
        plan.configure(nt, stage, 1.0)
        desc = plan.describe()
        say(desc)
        assert f"nt={nt} " in desc and f"stage_j={int(stage)}" in desc
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
    assert np.array_equal(dJ.cpu().numpy(), want)
    # the same configuration through the batch path (channel table of the F channels)
    specs = [(T, c) for c in range(F)] + [(77, 1)]
    host, dev = make_chunks(specs, D, B, 600)
    with bf.Plan(bf.Shape.make(1, F, D, B), A, G) as table:
        table.configure(nt, stage, 1.0)
        with bf.Batch(table, 8) as batch:
            batch.run(chunk_list(specs, dev), stream())
            torch.cuda.synchronize()
            say(batch.describe())
    check_chunks(specs, host, dev, A, G, D, B)


def test_configure_rejects_four_beam_tiles():
    # at most 3 beam tiles per workgroup (registers); the plan keeps working
    T, F, D, B = 64, 1, 512, 128
    A = bo.fill_uniform(F * B * D, 25)
    G = bo.setup_gain(F * B, D)
    with bf.Plan(bf.Shape.make(T, F, D, B), A, G) as plan:
        with pytest.raises(bf.BeamformError):
            plan.configure(4, False, 1.0)
        plan.configure(2, True, 1.0)
        assert "nt=2 " in plan.describe()


def test_batch_follows_plan_reconfiguration():
    # A batch made before its plan is reconfigured (other beam tiling, LDS
    # layout, workgroup count, new weight tables) or given new weights keeps
    # giving exact results, through the captured graph too; gains that move
    # the plan off the MFMA path make batch runs fail cleanly.
    C, D, B = 4, 512, 96
    A, G = channel_table(C, D, B, 31)
    specs = [(900, 0), (333, 3), (2048, 1), (77, 2)]
    host, dev = make_chunks(specs, D, B, 700)

    def run_and_check(batch, graph, A, G):
        for _, dJ in dev:
            dJ.zero_()
        torch.cuda.synchronize()
        batch.run(chunk_list(specs, dev), stream(), graph=graph)
        torch.cuda.synchronize()
        say(f"  {batch.describe()}")
        check_chunks(specs, host, dev, A, G, D, B)

    with bf.Plan(bf.Shape.make(1, C, D, B), A, G) as table, bf.Batch(table, 8) as batch:
        run_and_check(batch, True, A, G)
        for nt, stage, scale in [(1, False, 1.0), (2, False, 2.0), (3, True, 0.5)]:
            table.configure(nt, stage, scale)
            say(table.describe())
            run_and_check(batch, True, A, G)
            run_and_check(batch, False, A, G)
        A2, G2 = channel_table(C, D, B, 32)
        table.update_weights(A2, G2)
        run_and_check(batch, True, A2, G2)
        Gbig = np.full_like(G, 1e30)  # outside the fast requantisation range: generic kernel
        table.update_weights(A2, Gbig)
        with pytest.raises(bf.BeamformError) as ei:
            batch.run(chunk_list(specs, dev), stream())
        assert ei.value.status == bf.BF_ERR_UNSUPPORTED
        torch.cuda.synchronize()
        table.update_weights(A, G)
        run_and_check(batch, True, A, G)


@pytest.mark.parametrize("layout", [bf.BF_J_BFPT, bf.BF_J_BFTP])
def test_multi_channel_plain_and_batched(layout):
    C, D, B = 20, 512, 96
    A, G = channel_table(C, D, B, 15)
    shape = bf.Shape.make(1, C, D, B, j_layout=layout)
    specs = [(1000, 3, 4), (17, 0, 16), (2500, 19, 1), (64, 4, 16)]  # (ntimes, channel, nfreq)
    host, dEs, dJp, dJb = [], [], [], []
    for i, (T, c, nf) in enumerate(specs):
        E = bo.fill_uniform(T * nf * D * 2, 600 + i)
        host.append(E)
        dEs.append(torch.from_numpy(E).cuda())
        dJp.append(torch.zeros(B * nf * 2 * T, dtype=torch.uint8, device="cuda"))
        dJb.append(torch.zeros(B * nf * 2 * T, dtype=torch.uint8, device="cuda"))
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 8) as batch:
        for (T, c, nf), dE, dJ in zip(specs, dEs, dJp):
            plan.form_beams_channels(c, nf, T, dJ.data_ptr(), dE.data_ptr(), stream())
        batch.run([(dE.data_ptr(), dJ.data_ptr(), T, c, nf) for (T, c, nf), dE, dJ in zip(specs, dEs, dJb)], stream())
        torch.cuda.synchronize()
    for i, ((T, c, nf), E) in enumerate(zip(specs, host)):
        want = bo.form_beams(E, A.reshape(C, -1)[c:c + nf], G.reshape(C, -1)[c:c + nf], T, nf, D, B, jlayout=layout)
        assert np.array_equal(dJp[i].cpu().numpy(), want), f"plain launch, chunk {i}"
        assert np.array_equal(dJb[i].cpu().numpy(), want), f"batched, chunk {i}"


@pytest.mark.parametrize("D,B,layout", [(512, 96, bf.BF_J_BFPT), (512, 96, bf.BF_J_BFTP), (256, 192, bf.BF_J_BFPT),
                                        (520, 128, bf.BF_J_BFPT)])
def test_small_launches_with_one_beam_tile_per_workgroup(D, B, layout):
    # launches that would leave CUs idle with the plan's beam tiling run one
    # beam tile per workgroup over the same weight tables: exact as well
    C = 6
    A, G = channel_table(C, D, B, 41)
    shape = bf.Shape.make(1, C, D, B, j_layout=layout)
    specs = [(1024, 0), (16, 5), (3000, 2), (16384, 1), (1, 3)]
    host, dev = make_chunks(specs, D, B, 900)
    with bf.Plan(shape, A, G) as plan:
        desc = plan.describe()
        say(f"D={D} B={B} layout={layout}: {desc}")
        assert "launches under one workgroup per CU: nt=1" in desc
        for (T, c), (dE, dJ) in zip(specs, dev):
            plan.form_beams_channel(c, T, dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
        check_chunks(specs, host, dev, A, G, D, B, layout)
        # several channels in one launch
        T, c0, nf = 700, 1, 4
        E = bo.fill_uniform(T * nf * D * 2, 77)
        dE = torch.from_numpy(E).cuda()
        dJ = torch.zeros(B * nf * 2 * T, dtype=torch.uint8, device="cuda")
        plan.form_beams_channels(c0, nf, T, dJ.data_ptr(), dE.data_ptr(), stream())
        torch.cuda.synchronize()
    Ac = A.reshape(C, B * D)[c0:c0 + nf].reshape(-1)
    Gc = G.reshape(C, B)[c0:c0 + nf].reshape(-1)
    assert np.array_equal(dJ.cpu().numpy(), bo.form_beams(E, Ac, Gc, T, nf, D, B, jlayout=layout))
    # a whole small plan (its own shape) takes the same route
    T, F = 300, 2
    shape2 = bf.Shape.make(T, F, D, B, j_layout=layout)
    E2 = bo.fill_uniform(T * F * D * 2, 78)
    with bf.Plan(shape2, A[:F * B * D], G[:F * B]) as plan2:
        dE2 = torch.from_numpy(E2).cuda()
        dJ2 = torch.zeros(shape2.j_bytes, dtype=torch.uint8, device="cuda")
        plan2.form_beams_device(dJ2.data_ptr(), dE2.data_ptr(), stream())
        torch.cuda.synchronize()
    assert np.array_equal(dJ2.cpu().numpy(), bo.form_beams(E2, A[:F * B * D], G[:F * B], T, F, D, B, jlayout=layout))


def test_small_batches_take_one_beam_tile_per_workgroup():
    # a balanced batch too small to give every workgroup an item runs one beam
    # tile per workgroup; alternating small and large batches (eager and
    # graph replay, whose graphs are captured again when the choice changes)
    # stay exact
    C, D, B = 6, 512, 96
    A, G = channel_table(C, D, B, 43)
    small = [(2048, 0), (700, 3), (16, 5)]
    large = [(20000, c % C) for c in range(12)]
    hs, ds = make_chunks(small, D, B, 1100)
    hl, dl = make_chunks(large, D, B, 1200)
    with bf.Plan(bf.Shape.make(1, C, D, B), A, G) as plan, bf.Batch(plan, 16) as batch:
        for graph in (False, True, True, False, True):
            for specs, host, dev, want_nt in ((small, hs, ds, 1), (large, hl, dl, 3)):
                for _, dJ in dev:
                    dJ.zero_()
                torch.cuda.synchronize()
                batch.run(chunk_list(specs, dev), stream(), graph=graph)
                torch.cuda.synchronize()
                desc = batch.describe()
                say(f"  graph={graph}: {desc}")
                assert f"nt={want_nt} " in desc
                check_chunks(specs, host, dev, A, G, D, B)


def test_batch_unaligned_chunks():
    # chunks whose E and J start at odd byte addresses, plain and graph runs
    C, D, B = 4, 512, 96
    A, G = channel_table(C, D, B, 15)
    shape = bf.Shape.make(1, C, D, B)
    specs = [(1000, 0, 1, 3), (2049, 3, 5, 1), (16, 1, 0, 7), (3333, 2, 9, 2)]  # (T, channel, E offset, J offset)
    host, bufs = [], []
    for i, (T, c, eo, jo) in enumerate(specs):
        E = bo.fill_uniform(T * D * 2, 700 + i)
        dE = torch.zeros(T * D * 2 + 32, dtype=torch.uint8, device="cuda")
        dE[eo:eo + E.size] = torch.from_numpy(E).cuda()
        host.append(E)
        bufs.append((dE, torch.zeros(B * 2 * T + 32, dtype=torch.uint8, device="cuda")))
    chunks = [(dE.data_ptr() + eo, dJ.data_ptr() + jo, T, c) for (T, c, eo, jo), (dE, dJ) in zip(specs, bufs)]
    with bf.Plan(shape, A, G) as plan, bf.Batch(plan, 8) as batch:
        for graph in (False, True):
            for _, dJ in bufs:
                dJ.zero_()
            torch.cuda.synchronize()
            batch.run(chunks, stream(), graph=graph)
            torch.cuda.synchronize()
            for i, ((T, c, eo, jo), E, (_, dJ)) in enumerate(zip(specs, host, bufs)):
                J = dJ.cpu().numpy()
                assert (J[:jo] == 0).all() and (J[jo + B * 2 * T:] == 0).all(), f"chunk {i}: bytes outside J written"
                assert np.array_equal(J[jo:jo + B * 2 * T], oracle_chunk(E, A, G, c, T, D, B)), f"chunk {i} graph={graph}"


def test_batch_items_of_many_segments():
    # thousands of short chunks: a balanced work item then spans far more
    # segments than a workgroup holds in LDS at a time (64), so the kernel
    # reads its segment records in several batches; every chunk stays exact
    C, D, B = 8, 64, 32
    A, G = channel_table(C, D, B, 51)
    rng = np.random.default_rng(52)
    n = 80000  # one round each: about 80 segments per item on up to 1024 workgroups
    Ts = rng.integers(1, 25, n)
    chans = rng.integers(0, C, n)
    e_off = np.concatenate([[0], np.cumsum(Ts * D * 2)])
    j_off = np.concatenate([[0], np.cumsum(Ts * B * 2)])
    E = bo.fill_uniform(int(e_off[-1]), 53)
    dE = torch.from_numpy(E).cuda()
    dJ = torch.zeros(int(j_off[-1]), dtype=torch.uint8, device="cuda")
    chunks = [(dE.data_ptr() + int(e_off[i]), dJ.data_ptr() + int(j_off[i]), int(Ts[i]), int(chans[i]))
              for i in range(n)]
    with bf.Plan(bf.Shape.make(1, C, D, B), A, G) as plan, bf.Batch(plan, n) as batch:
        batch.run(chunks, stream())
        torch.cuda.synchronize()
        desc = batch.describe()
        say(desc)
        items = int(desc.split(" items=")[1].split()[0])
        segments = int(desc.split(" segments=")[1].split()[0])
        assert segments > 64 * items, "few segments per item: the LDS batching is not exercised"
    J = dJ.cpu().numpy()
    for i in range(n):
        want = oracle_chunk(E[e_off[i]:e_off[i + 1]], A, G, int(chans[i]), int(Ts[i]), D, B)
        assert np.array_equal(J[j_off[i]:j_off[i + 1]], want), f"chunk {i} (T={Ts[i]}, channel={chans[i]})"
