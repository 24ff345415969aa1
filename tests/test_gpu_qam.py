"""The QAM half of the device trial generator against the oracle, through the C ABI.

- LLR values: trial by trial, the device's FP32 LLRs equal the oracle's restatement of the
  same generator (the same Philox counters, Box-Muller, encoder and Gray mapping, then the
  max-log LLRs of demapper.rs:270-297) in f64, to FP32 rounding; information bits exactly.
- Statistics: the device 16-QAM chain's FER on NR BG1 Z=384 (layered NMS 0.75, 10
  iterations) lies inside the 95 % Wilson interval (stats.rs:74-87) of the oracle's run of
  the reference trial chain (xoshiro streams of rng.rs, complex_gaussian noise, f64 decoder) on
  6000 frames.
"""
import math
import os

import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def noise_variance(code, mod, snr_kind, snr_db):
    """The generator's noise: BPSK sigma^2 per real dimension, QAM N0 per symbol (bounds.rs:58-60,
    awgn.rs:139-141 with unit-energy symbols)."""
    lin = 10.0 ** (snr_db / 10.0)
    rate = code.k / code.n
    if mod == F.BPSK:
        return 1.0 / (2.0 * rate * lin) if snr_kind == F.EBN0_DB else 1.0 / (2.0 * lin)
    return 1.0 / (mod * rate * lin) if snr_kind == F.EBN0_DB else 1.0 / lin


CASES = [("BG1 Z=384 16-QAM Es/N0 4 dB", (8448, 25344, 1, 0, 4), F.QAM16, F.ESN0_DB, 4.0),
         ("BG1 Z=384 QPSK Eb/N0 1 dB", (8448, 25344, 1, 0, 2), F.QPSK, F.EBN0_DB, 1.0),
         ("K'=1000 E=1500 rv3 64-QAM Es/N0 12 dB", (1000, 1500, 0, 3, 6), F.QAM64, F.ESN0_DB, 12.0),
         ("K'=500 E=1200 rv2 BPSK Eb/N0 2 dB", (500, 1200, 0, 2, 1), F.BPSK, F.EBN0_DB, 2.0)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_device_llrs_equal_the_oracle_restatement(case):
    name, args, mod, kind, snr = case
    print(f"\n[qam] device LLRs vs oracle: {name}", flush=True)
    code, ref = F.NrLdpc(*args), O.Code.nr(*args)
    llr, info = F.generate_batch(code, 4, mod, snr, snr_kind=kind, seed=13, point=2, trial0=77)
    torch.cuda.synchronize()
    llr = llr.cpu().numpy().astype(np.float64)
    bits = F.unpack_info(info, code.k)
    n0 = noise_variance(code, mod, kind, snr)
    worst = 0.0
    for t in range(4):
        want_info, _, want = ref.philox_trial(mod, n0, 13, 2, 77 + t)
        assert np.array_equal(bits[t], want_info)
        # FP32 rounding of the metrics (|metric| up to ~ (|y| + 1)^2 / N0) and of the noise
        err = np.abs(llr[t] - want)
        tol = 1e-4 + 1e-5 * np.abs(want)
        assert np.all(err <= tol), (t, float(err.max()), int(err.argmax()))
        worst = max(worst, float((err / (1.0 + np.abs(want))).max()))
        # the decisions the decoder starts from are the same
        assert np.array_equal(llr[t] < 0, want < 0) or np.all(np.abs(want[(llr[t] < 0) != (want < 0)]) < 1e-4)
    print(f"  largest |error| / (1 + |LLR|): {worst:.2e}", flush=True)


def device_fer(code, esn0, frames, seed):
    fe = 0
    for part in range(frames // 4096):
        llr, info = F.generate_batch(code, 4096, F.QAM16, esn0, snr_kind=F.ESN0_DB, seed=seed, point=0,
                                     trial0=4096 * part)
        bits, _, _ = code.decode_batch(llr)
        fe += int((bits.cpu().numpy() != F.unpack_info(info, code.k)).any(1).sum())
    return fe


@pytest.mark.parametrize("esn0", [3.9, 4.0])
def test_device_16qam_fer_inside_the_oracle_chain_wilson95(esn0):
    print(f"\n[qam] 16-QAM FER at Es/N0 {esn0} dB: device vs oracle chain ({THREADS} host threads)", flush=True)
    args = (8448, 25344, 1, 0, 4)
    code = F.NrLdpc(*args).set_decoder("normalized_min_sum", 0.75, True, 10)
    frames = 16384
    fe = device_fer(code, esn0, frames, seed=21)
    ref = O.Code.nr(*args)
    n0 = 10.0 ** (-esn0 / 10.0)
    n_ref = 6000
    info, llr = ref.trial_nr_qam_many(16, n0, 1, "qam16", 0, 0, n_ref, threads=THREADS)
    fulls = np.array([ref.full_llr(l) for l in llr])
    post, _ = ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=THREADS)
    fe_ref = int(((post[:, :ref.k_info] < 0).astype(np.uint8) != info).any(1).sum())
    lo, hi = O.wilson(fe_ref, n_ref, 0.95)
    print(f"  device {fe}/{frames} = {fe / frames:.4f}; oracle {fe_ref}/{n_ref} = {fe_ref / n_ref:.4f} "
          f"[{lo:.4f}, {hi:.4f}]", flush=True)
    assert lo <= fe / frames <= hi
