#!/usr/bin/env python3
"""Benchmark of the MI355X form-beams path (one GPU).

Default workload (the full-size single-shot configuration): T=32768 time
samples, F=32 frequencies, D=512 dishes, P=2 polarizations, B=96 beams;
E and A are seeded uniform ucomplex4 values in [-7, 7]^2, G is the
reference's gain formula. E is generated in HBM before timing; a "step" is
one form_beams launch over the whole chunk (J = 201 MB written).

Prints one JSON line (stdout) with:
  value          algorithmic int8 TOPS (8 T F B D P ops per step)
  roofline       dominant kernel vs the MI355X dense int8 peak
  cpu_baseline   the oracle (CPU restatement of the reference cpu.cxx)
                 on a bounded sample of the same workload on this host
  beam_samples_per_s, hbm_gbps, parity of the sampled frequencies

Other modes:
  --copy-calib BYTES   run only a plain device copy of BYTES bytes (used to
                       calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

MI355X_INT8_TOPS = 5000.0  # dense int8 MFMA, 5.0 POP/s
MI355X_HBM_GBPS = 8000.0   # HBM3E, 8 TB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload_bytes(T, F, D, B, P=2):
    return T * F * D * P + F * B * D + 4 * F * B + B * F * T * P


def workload_ops(T, F, D, B, P=2):
    return 8 * T * F * B * D * P


def load_traffic(path):
    """HBM bytes per launch measured with rocprofv3 (profiles/*.json), if any."""
    try:
        with open(path) as fh:
            return json.load(fh).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(E_host, A, G, T, F, D, B, f_sample, J_gpu):
    """Times the oracle on frequencies [0, f_sample) of the same inputs and
    checks those outputs against the GPU's."""
    from oracle import bf_oracle as bo
    Es = np.ascontiguousarray(E_host.reshape(T, F, D * 2)[:, :f_sample])
    As = np.ascontiguousarray(A.reshape(F, B * D)[:f_sample])
    Gs = np.ascontiguousarray(G.reshape(F, B)[:f_sample])
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    Jc = bo.form_beams(Es, As, Gs, T, f_sample, D, B, nthreads=threads)
    dt = time.perf_counter() - t0
    ops = workload_ops(T, f_sample, D, B)
    Jg = J_gpu.reshape(B, F, 2, T)[:, :f_sample].reshape(-1)
    parity = bool(np.array_equal(Jc, Jg))
    return {
        "value": ops / dt / 1e12,
        "unit": "TOPS (int8, algorithmic)",
        "cores": min(threads, B),
        "threads_requested": threads,
        "kind": "port",
        "sample": f"oracle form_beams on T={T} F={f_sample} (of {F}) D={D} P=2 B={B}, same inputs; "
                  f"{dt:.2f} s wall; OpenMP over beams (parallelism capped at B={B})",
        "seconds": dt,
        "beam_samples_per_s": B * f_sample * T * 2 / dt,
        "parity_vs_gpu": parity,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ntimes", type=int, default=32768)
    ap.add_argument("--nfrequencies", type=int, default=32)
    ap.add_argument("--ndishes", type=int, default=512)
    ap.add_argument("--nbeams", type=int, default=96)
    ap.add_argument("--cpu-frequencies", type=int, default=4, help="frequencies in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calib", type=int, default=0, help="only run a device copy of this many bytes")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the form-beams path shards by frequency with no exchange "
              "between GPUs, so one GPU's throughput is the figure (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    import beamforming_amd as bf
    from oracle import bf_oracle as bo

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    if args.copy_calib:
        src = torch.empty(args.copy_calib, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calib} bytes x5")
        return

    T, F, D, B = args.ntimes, args.nfrequencies, args.ndishes, args.nbeams
    shape = bf.Shape.make(T, F, D, B)
    A = bo.fill_uniform(F * B * D, 12)
    G = bo.setup_gain(F * B, D)
    dE = torch.empty(shape.e_bytes, dtype=torch.uint8, device=dev)
    bf.fill_uniform_device(dE.data_ptr(), shape.e_bytes, 11, 0, sptr)
    dJ = torch.empty(shape.j_bytes, dtype=torch.uint8, device=dev)
    plan = bf.Plan(shape, A, G)
    log(f"kernel: {plan.describe()}")

    for _ in range(args.warmup):
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), sptr)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.form_beams_device(dJ.data_ptr(), dE.data_ptr(), sptr)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps

    ops = workload_ops(T, F, D, B)
    nbytes = workload_bytes(T, F, D, B)
    tops = ops / (ms * 1e-3) / 1e12
    gbps = nbytes / (ms * 1e-3) / 1e9
    checksum = bf.adler32_device(dJ.data_ptr(), shape.j_bytes, sptr)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(dE.cpu().numpy(), A, G, T, F, D, B, min(args.cpu_frequencies, F), dJ.cpu().numpy())
        log(f"cpu: {cpu['value']:.4f} TOPS in {cpu['seconds']:.2f} s, parity={cpu['parity_vs_gpu']}")

    traffic = load_traffic(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                        "form_beams_full_traffic.json"))
    out = {
        "metric": "int8 TOPS & beam-samples/s (% roofline) vs plain launch; p99 chunk latency",
        "value": round(tops, 2),
        "unit": "TOPS (int8, algorithmic 8*T*F*B*D*P ops)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (seeded uniform ucomplex4 E and A in [-7,7]^2, reference gain formula)",
        "config": {"workload": "full-size single shot", "ntimes": T, "nfrequencies": F, "ndishes": D,
                   "npolarizations": 2, "nbeams": B, "j_layout": "[b][f][p][t]"},
        "roofline": {"bound": "mfma", "achieved": round(tops, 2), "peak": MI355X_INT8_TOPS, "unit": "TFLOP/s",
                     "frac": round(tops / MI355X_INT8_TOPS, 4), "traffic": traffic,
                     "note": "int8 MFMA ops per second vs the 5.0 POP/s dense int8 peak; achieved = algorithmic "
                             "ops per launch / mean launch time (HIP events)"},
        "beam_samples_per_s": B * F * T * 2 / (ms * 1e-3),
        "hbm_gbps_algorithmic": round(gbps, 1),
        "hbm_frac": round(gbps / MI355X_HBM_GBPS, 4),
        "kernel": plan.describe(),
        "j_adler32": f"0x{checksum:08x}",
        "wall_s_timed_region": round(wall, 4),
        "cpu_baseline": cpu,
    }
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
