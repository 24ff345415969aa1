"""The BG1 QPSK campaign: frame and bit error rates of NR base graph 1, Z=384 (K'=8448,
E=25344, rate 1/3, QPSK with the NR bit interleaver) over AWGN at Eb/N0 = 0, 0.25, ...,
1.5 dB, plain layered normalized min-sum (alpha 0.75, 10 iterations) against the learned
offset/scale min-sum retrained for this link, each point run by the Runner until 200 frame
errors or 10^6 trials (Stop::min_errors("fer", 200) or Stop::max_trials(1_000_000),
crates/com-sim/src/stopping.rs), trials keyed by Philox (seed 7, point, trial).

    python com-rs_amd/campaign_bg1_qpsk.py [--max-trials N] [--learned-iterations T]

Prints one JSON line: per point and decoder the trials, frame errors, FER with its 95 %
Wilson interval, BER, mean iterations and the wall time of the point.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

POINTS = [0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5]


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--max-trials", type=int, default=1_000_000)
    p.add_argument("--min-errors", type=int, default=200)
    p.add_argument("--batch", type=int, default=16384)
    p.add_argument("--learned-iterations", type=int, default=8)
    p.add_argument("--train-steps", type=int, default=1000)
    p.add_argument("--fallback", type=int, default=1)
    a = p.parse_args()
    import torch
    import comfec as F

    torch.cuda.init()
    code = F.NrLdpc(8448, 25344, 1, 0, 2).set_decoder("normalized_min_sum", 0.75, True, 10)
    grid = F.Grid().axis("ebn0_db", POINTS)
    stop = F.Stop.min_errors("fer", a.min_errors) | F.Stop.max_trials(a.max_trials)

    def run(params):
        t0 = time.perf_counter()
        res = F.Runner().seed(7).batch(a.batch).stop(stop).run_nr(
            code, grid, F.QPSK, F.EBN0_DB, learned_params=params, fallback=bool(a.fallback))
        wall = time.perf_counter() - t0
        rows = []
        for r in res:
            fer, ber, its = r.metric("fer"), r.metric("ber"), r.metric("iterations")
            rows.append({"ebn0_db": r.condition["ebn0_db"], "trials": r.trials, "frame_errors": fer.count,
                         "fer": fer.value, "fer_wilson95": [fer.lo, fer.hi], "ber": ber.value,
                         "mean_iterations": its.value, "wall_s": round(r.wall_s, 3)})
        return rows, wall

    plain, plain_wall = run(None)
    params = F.initial_params(code, a.learned_iterations)
    t0 = time.perf_counter()
    hist = F.learned_train(code, params, steps=a.train_steps, batch=512, lr=1e-2, modulation=F.QPSK,
                           snr_kind=F.EBN0_DB, snr_lo=0.5, snr_hi=1.5, seed=11)
    train_s = time.perf_counter() - t0
    learned, learned_wall = run(params)
    print(json.dumps({
        "workload": "NR BG1 Z=384 K'=8448 E=25344 QPSK (NR interleaver), AWGN, Eb/N0 0..1.5 dB, "
                    f"min_errors fer {a.min_errors} or max_trials {a.max_trials}, seed 7",
        "plain": {"decoder": "layered NMS alpha 0.75, 10 it", "points": plain, "wall_s": round(plain_wall, 2)},
        "learned": {"decoder": f"offset/scale min-sum, {a.learned_iterations} it"
                               + (", plain fallback" if a.fallback else ""),
                    "train_steps": a.train_steps, "train_s": round(train_s, 2),
                    "loss_first_last": [float(hist[:20].mean()), float(hist[-20:].mean())],
                    "points": learned, "wall_s": round(learned_wall, 2)},
    }), flush=True)


if __name__ == "__main__":
    main()
