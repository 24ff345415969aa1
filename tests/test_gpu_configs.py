"""The two statistical workloads end to end on the GPU, through the C ABI.

- The BG1 QPSK campaign: NR base graph 1, Z=384 (K'=8448, E=25344, rate 1/3, QPSK with the
  NR bit interleaver) over AWGN at Eb/N0 = 0, 0.25, ..., 1.5 dB, each point run by the
  Runner until Stop::min_errors("fer", 200) | Stop::max_trials(1_000_000)
  (crates/com-sim/src/stopping.rs:94-141), trials keyed by Philox (seed 7, point, trial);
  the plain layered normalized min-sum (alpha 0.75, 10 iterations) against the learned
  offset/scale min-sum retrained on this link, with the plain fallback.
- The learned decoder on BG2 Z=52 (K=520, 2600 bits sent): 12 unrolled layered iterations,
  42 x 12 x 2 FP32 parameters trained with Adam at lr 1e-3 on batches of 8192 generator
  trials (BPSK, Eb/N0 uniform in [0, 3] dB, binary cross-entropy on the soft outputs),
  evaluated against the plain normalized min-sum and the f64 sum-product rule.
"""
import json

import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

POINTS = [0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5]


def campaign_rows(res):
    rows = []
    for r in res:
        fer, ber, its = r.metric("fer"), r.metric("ber"), r.metric("iterations")
        rows.append({"ebn0_db": r.condition["ebn0_db"], "status": r.status, "trials": r.trials,
                     "frame_errors": fer.count, "fer": fer.value, "lo": fer.lo, "hi": fer.hi, "ber": ber.value,
                     "mean_iterations": round(its.value, 3)})
    return rows


class BatchLog:
    """A callback that keeps (trials, frame errors) of every condition after every batch."""

    def __init__(self):
        self.seen = {}

    def on_batch_end(self, progress):
        self.seen.setdefault(progress.condition, []).append((progress.trials, progress.metric("fer").count))


def test_bg1_qpsk_campaign_learned_vs_plain():
    print("\n[config 5] BG1 Z=384 QPSK campaign, Eb/N0 0..1.5 dB, 200 frame errors or 1e6 trials", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 2).set_decoder("normalized_min_sum", 0.75, True, 10)
    grid = F.Grid().axis("ebn0_db", POINTS)
    batch = 16384
    stop = F.Stop.min_errors("fer", 200) | F.Stop.max_trials(1_000_000)
    logs = [BatchLog(), BatchLog()]
    plain = campaign_rows(F.Runner().seed(7).batch(batch).stop(stop).callback(logs[0]).run_nr(
        code, grid, F.QPSK, F.EBN0_DB))
    params = F.initial_params(code, 8)
    hist = F.learned_train(code, params, steps=1000, batch=512, lr=1e-2, modulation=F.QPSK, snr_kind=F.EBN0_DB,
                           snr_lo=0.5, snr_hi=1.5, seed=11)
    assert np.all(np.isfinite(hist)) and hist[-50:].mean() < hist[:50].mean()
    learned = campaign_rows(F.Runner().seed(7).batch(batch).stop(stop).callback(logs[1]).run_nr(
        code, grid, F.QPSK, F.EBN0_DB, learned_params=params, fallback=True))
    print(json.dumps({"plain": plain, "learned": learned}), flush=True)
    for rows, log in zip((plain, learned), logs):
        assert [r["ebn0_db"] for r in rows] == POINTS
        for c, r in enumerate(rows):
            # the stop (runner.rs:429-470): batches of `batch` trials, the last cut at the trial
            # limit; is_met after every batch, so every batch but the last left the rule unmet
            steps = log.seen[c]
            assert [t for t, _ in steps] == [min(batch * (i + 1), 1_000_000) for i in range(len(steps))]
            assert steps[-1] == (r["trials"], r["frame_errors"])
            assert all(fe < 200 and t < 1_000_000 for t, fe in steps[:-1])
            assert r["status"] == "completed"
            assert r["frame_errors"] >= 200 or r["trials"] == 1_000_000
            lo, hi = O.wilson(r["frame_errors"], r["trials"], 0.95)
            assert (r["lo"], r["hi"]) == pytest.approx((lo, hi), rel=1e-12, abs=1e-15)
        # the FER falls along the grid
        fers = [r["fer"] for r in rows]
        assert all(a >= b for a, b in zip(fers, fers[1:]))
    p, l = plain[5], learned[5]
    assert p["ebn0_db"] == 1.25 and l["hi"] < p["lo"], (p, l)
    assert plain[6]["frame_errors"] == 0 and plain[6]["hi"] < 1e-5
    assert learned[6]["frame_errors"] == 0 and learned[6]["hi"] < 1e-5


def test_learned_bg2_z52_recipe_against_plain_and_sum_product():
    print("\n[config 4] learned BG2 Z=52, 12 it, Adam lr 1e-3 batch 8192", flush=True)
    T = 12
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, T)
    spa = F.NrLdpc.lifted(2, 52).set_decoder("sum_product", 0.0, True, T)
    params = F.initial_params(code, T)
    hist = F.learned_train(code, params, steps=1500, batch=8192, lr=1e-3, modulation=F.BPSK, snr_kind=F.EBN0_DB,
                           snr_lo=0.0, snr_hi=3.0, seed=4)
    assert np.all(np.isfinite(hist))
    print(f"  loss {hist[:100].mean():.5f} -> {hist[-100:].mean():.5f}", flush=True)
    assert hist[-100:].mean() < hist[:100].mean() - 0.005
    out = {}
    for pi, ebn0 in enumerate((1.0, 1.5)):
        n = 65536
        acc = {k: [0, 0, 0.0] for k in ("plain", "learned", "learned_fallback", "spa_f64")}
        for s in range(0, n, 16384):
            llr, info = F.generate_batch(code, 16384, F.BPSK, ebn0, snr_kind=F.EBN0_DB, seed=21, point=pi, trial0=s)
            truth = F.unpack_info(info, code.k)
            runs = {"plain": code.decode_batch(llr)[:2],
                    "learned": F.learned_decode_batch(code, params, llr, fallback=False)[:2],
                    "learned_fallback": F.learned_decode_batch(code, params, llr, fallback=True)[:2]}
            if s < 32768:
                runs["spa_f64"] = spa.decode_batch(llr.double())[:2]
            for k, (bits, its) in runs.items():
                acc[k][0] += int((bits.cpu().numpy() != truth).any(1).sum())
                acc[k][1] += 16384
                acc[k][2] += float(its.double().sum().item())
        out[ebn0] = {k: {"frames": nn, "frame_errors": fe, "fer": fe / nn, "wilson95": F.wilson(fe, nn, 0.95),
                         "mean_iterations": round(it / nn, 3)} for k, (fe, nn, it) in acc.items()}
    print(json.dumps(out), flush=True)
    for ebn0 in (1.0, 1.5):
        r = out[ebn0]
        # the trained decoder beats the plain rule; the f64 sum-product judge beats both
        assert r["learned"]["wilson95"][1] < r["plain"]["wilson95"][0]
        assert r["spa_f64"]["wilson95"][1] < r["learned"]["wilson95"][0]
        # the fallback re-decodes with the plain rule only the frames whose syndrome failed
        assert r["learned_fallback"]["mean_iterations"] >= r["learned"]["mean_iterations"]
