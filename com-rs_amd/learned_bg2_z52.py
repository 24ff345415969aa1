"""The learned decoder on NR base graph 2, Z=52 (K=520, 2600 bits sent, the first 2Z not
sent): 12 unrolled layered iterations with one offset and one scale per base row and
iteration (42 x 12 x 2 FP32 parameters), trained on the GPU with Adam (lr 1e-3, batches of
8192 trials, binary cross-entropy on the soft outputs) on BPSK/AWGN trials of the device
generator at Eb/N0 uniform in [0, 3] dB; then the frame error rates, on fresh trials, of
the learned decoder (alone and with the plain fallback), the plain layered normalized
min-sum (alpha 0.75) and the sum-product rule in f64, all with 12 iterations.

    python com-rs_amd/learned_bg2_z52.py [--train-steps N] [--frames N] [--spa-frames N]

Prints the training progress every 1000 steps and each point's results on stderr, then
one JSON line: the loss curve (means of 100-step windows) and, per Eb/N0, each decoder's
frames, frame errors, FER with its 95 % Wilson interval and mean iterations.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

POINTS = [0.5, 1.0, 1.5, 2.0, 2.5]
ITERATIONS = 12


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--train-steps", type=int, default=16000)
    p.add_argument("--batch", type=int, default=8192)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--frames", type=int, default=131072)
    p.add_argument("--spa-frames", type=int, default=32768)
    a = p.parse_args()
    import numpy as np
    import torch
    import comfec as F

    torch.cuda.init()
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, ITERATIONS)
    spa = F.NrLdpc.lifted(2, 52).set_decoder("sum_product", 0.0, True, ITERATIONS)
    params = F.initial_params(code, ITERATIONS)
    # the steps of cf_learned_train one by one (fresh trials (seed 4, point 0) every step,
    # one Adam state throughout), so that the run reports its progress
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    losses = []
    t0 = time.perf_counter()
    for step in range(1, a.train_steps + 1):
        llr, cw = F.generate_training_batch(code, a.batch, F.BPSK, 0.0, 3.0, snr_kind=F.EBN0_DB, seed=4,
                                            trial0=(step - 1) * a.batch)
        loss, grad = F.learned_loss_grad(code, params, llr, cw)
        F.adam_step(params, grad, m, v, step, lr=a.lr)
        losses.append(loss)
        if step % 1000 == 0:
            recent = float(torch.cat(losses[-100:]).mean().item())
            print(json.dumps({"step": step, "loss_last_100": round(recent, 5),
                              "seconds": round(time.perf_counter() - t0, 1)}), file=sys.stderr, flush=True)
    hist = torch.cat(losses).cpu().numpy() if losses else np.zeros(0)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    windows = [round(float(hist[i:i + 100].mean()), 5) for i in range(0, len(hist), 100)]

    def wil(fe, n):
        return [round(x, 8) for x in F.wilson(fe, n, 0.95)]

    points = []
    chunk = 16384
    for pi, ebn0 in enumerate(POINTS):
        acc = {k: [0, 0, 0.0] for k in ("plain", "learned", "learned_fallback", "spa_f64")}
        for s in range(0, a.frames, chunk):
            n = min(chunk, a.frames - s)
            llr, info = F.generate_batch(code, n, F.BPSK, ebn0, snr_kind=F.EBN0_DB, seed=21, point=pi, trial0=s)
            truth = F.unpack_info(info, code.k)
            runs = {"plain": code.decode_batch(llr)[:2],
                    "learned": F.learned_decode_batch(code, params, llr, fallback=False)[:2],
                    "learned_fallback": F.learned_decode_batch(code, params, llr, fallback=True)[:2]}
            if s < a.spa_frames:
                runs["spa_f64"] = spa.decode_batch(llr.double())[:2]
            for k, (bits, its) in runs.items():
                acc[k][0] += int((bits.cpu().numpy() != truth).any(1).sum())
                acc[k][1] += n
                acc[k][2] += float(its.double().sum().item())
        row = {"ebn0_db": ebn0}
        for k, (fe, n, it) in acc.items():
            row[k] = {"frames": n, "frame_errors": fe, "fer": fe / n, "fer_wilson95": wil(fe, n),
                      "mean_iterations": round(it / n, 3)}
        points.append(row)
        print(json.dumps({"progress": row}), file=sys.stderr, flush=True)
    print(json.dumps({
        "workload": "NR BG2 Z=52 lifted (K=520, 2600 sent), BPSK over AWGN, layered, 12 iterations",
        "training": {"steps": a.train_steps, "batch": a.batch, "lr": a.lr, "ebn0_db": [0.0, 3.0],
                     "seconds": round(train_s, 2), "loss_per_100_steps": windows},
        "points": points,
    }), flush=True)


if __name__ == "__main__":
    main()
