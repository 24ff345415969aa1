"""Throughput of the NR LDPC decoder on one MI355X.

Workload: NR base graph 1, Z=384 (K'=8448, E=25344, rate 1/3), Gray 16-QAM over AWGN,
max-log LLRs, layered normalized min-sum (alpha 0.75), 10 iterations with syndrome early
stop, FP32 messages, 4096 codewords per step. The Es/N0 of the timed steps is the point of
the grid {4, 5, 6, 7} dB whose FER is nearest 1e-3 (measured on the first batch).
A step decodes one batch of 4096 codewords whose LLRs are already resident in HBM.

Prints one JSON line (see the contract in the repository root docs).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "com-rs_amd"))

HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--iterations", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    return p.parse_args()


def cpu_baseline(code_args, llr_f32, seconds):
    """The oracle's f64 decoder (a C restatement of LdpcCode::propagate) on host threads,
    over a bounded sample of the same LLRs: decoded Gbit/s."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.Code.nr(*code_args)
    threads = max(1, min(16, os.cpu_count() or 1))
    # a first small run sizes the sample to about `seconds` of CPU work
    probe = min(threads, llr_f32.shape[0])
    fulls = np.array([ref.full_llr(l.astype(np.float64)) for l in llr_f32[:probe]])
    t0 = time.perf_counter()
    ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=threads)
    dt = time.perf_counter() - t0
    count = int(max(probe, min(llr_f32.shape[0], probe * seconds / max(dt, 1e-3))))
    fulls = np.array([ref.full_llr(l.astype(np.float64)) for l in llr_f32[:count]])
    t0 = time.perf_counter()
    ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": ref.k_info * count / dt / 1e9, "unit": "Gbit/s", "cores": threads, "kind": "port",
            "sample": f"{count} codewords of the timed batch (f64 layered NMS 0.75, 10 it, "
                      f"oracle C restatement, {threads} threads, {dt:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the decoder shards by "
              "codeword; see the multi-GPU paragraph of the design notes)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import comfec as F
    import workloads as W

    torch.cuda.init()
    dev = torch.device("cuda:0")
    k_info, e = 8448, 25344
    code = F.NrLdpc(k_info, e, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, args.iterations)
    batch = args.batch

    # calibration: FER at every grid point, on the batch that the timed steps will use
    points = {}
    for esn0 in (4.0, 5.0, 6.0, 7.0):
        info, llr = W.nr_qam_batch(code, batch, esn0, m=16, seed=1)
        bits, its, _ = code.decode_batch(llr)
        torch.cuda.synchronize()
        errs = (bits.cpu().numpy() != info).sum(1)
        fe = int((errs > 0).sum())
        points[esn0] = {"fer": fe / batch, "frame_errors": fe, "bit_errors": int(errs.sum()),
                        "mean_iterations": float(its.float().mean().item()), "llr": llr, "info": info}
    # the point nearest FER 1e-3 on a log scale (0 errors counts as 1e-5)
    esn0 = min(points, key=lambda p: abs(math.log10(max(points[p]["fer"], 1e-5)) + 3.0))
    llr = points[esn0]["llr"]
    for p in points:
        if p != esn0:
            points[p].pop("llr")
    mean_it = points[esn0]["mean_iterations"]

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        code.decode_batch(llr)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        code.decode_batch(llr)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)
    ms_per_step = ms_total / args.steps
    value = k_info * batch / (ms_per_step / 1e3) / 1e9

    # roofline of the decode kernel: algorithmic bytes per codeword (channel LLRs in, K/8
    # decisions out, and per iteration every edge message read and written in FP32)
    edges = 316 * 384
    alg_bytes_cw = e * 4 + math.ceil(k_info / 8) + mean_it * 2 * edges * 4
    achieved = alg_bytes_cw * batch / (ms_per_step / 1e3) / 1e9
    lds_bytes_cw = mean_it * edges * 16
    out = {
        "metric": "decoded Gbit/s at FER 1e-3, learned vs plain min-sum, NR BG1 Z=384",
        "value": round(value, 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: random info bits, NR encoder, Gray 16-QAM, AWGN, max-log LLRs (on device)",
        "config": {
            "workload": "plain layered NMS a=0.75, NR BG1 Z=384 K=8448 E=25344 rate 1/3, 16-QAM, 10 it",
            "decoder": "plain", "global_batch": batch, "esn0_db": esn0, "mean_iterations": round(mean_it, 3),
            "fer_by_esn0": {str(p): v["fer"] for p, v in points.items()},
            "wall_s_timed": round(wall, 4),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "lds": {"achieved": round(lds_bytes_cw * batch / (ms_per_step / 1e3) / 1e9, 2),
                             "peak": LDS_PEAK_GBS, "unit": "GB/s"}},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline((k_info, e, 1, 0, 4), llr.cpu().numpy(), args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
