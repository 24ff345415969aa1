"""The device trial generator and the GPU campaign, through the C ABI.

- information bits: bit-exact with the oracle's Philox-4x32-10 (pinned by its known answers);
- encoder, rate matching, interleaver and modulation: at a very high SNR the LLR signs are
  the oracle's encoded bits exactly;
- noise: LLR moments against the closed form of BPSK over AWGN;
- campaigns: the frame error rate inside the 95 % Wilson interval of the oracle's run of the
  reference trial chain (xoshiro streams), the bit error rate inside the 95 % interval of
  the mean of the oracle's per-frame bit-error fractions (bit errors come in bursts within a
  frame, so the frame is the independent unit), identical records whatever the batch size.
"""
import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_info_bits_are_the_philox_stream():
    print("\n[generator] info bits vs Philox", flush=True)
    code = F.NrLdpc.lifted(2, 52)
    llr, info = F.generate_batch(code, 32, F.BPSK, 2.0, seed=7, point=3, trial0=100)
    bits = F.unpack_info(info, code.k)
    for i in range(32):
        assert np.array_equal(bits[i], O.philox_info_bits(7, 3, 100 + i, code.k))
    # a trial's draws do not depend on the batch it is generated in
    llr2, info2 = F.generate_batch(code, 8, F.BPSK, 2.0, seed=7, point=3, trial0=110)
    assert torch.equal(info2, info[10:18]) and torch.equal(llr2, llr[10:18])


@pytest.mark.parametrize("case", [("lifted BG2 Z=52 BPSK", None, F.BPSK), ("NR BG1 Z=384 16-QAM qm=4", (8448, 25344, 1, 0, 4),
                                   F.QAM16), ("NR BG1 Z=384 QPSK qm=2", (8448, 25344, 1, 0, 2), F.QPSK),
                                  ("NR K'=500 E=1200 rv2 16-QAM", (500, 1200, 0, 2, 4), F.QAM16),
                                  ("NR K'=1000 E=1500 rv3 64-QAM", (1000, 1500, 0, 3, 6), F.QAM64)],
                         ids=lambda c: c[0])
def test_noiseless_llr_signs_are_the_encoded_bits(case):
    name, args, mod = case
    print(f"\n[generator] {name}", flush=True)
    if args is None:
        code, ref = F.NrLdpc.lifted(2, 52), O.Code.lifted(2, 52)
    else:
        code, ref = F.NrLdpc(*args), O.Code.nr(args[0], args[1], args[2], args[3], args[4])
    llr, info = F.generate_batch(code, 6, mod, 60.0, snr_kind=F.ESN0_DB, seed=11, point=0, trial0=0)
    llr = llr.cpu().numpy()
    bits = F.unpack_info(info, code.k)
    for i in range(6):
        if args is None:
            tx = ref.encode(bits[i])[2 * ref.z:]
        else:
            tx = ref.nr_encode(bits[i])
        assert np.array_equal((llr[i] < 0).astype(np.uint8), tx)
        assert np.all(np.abs(llr[i]) > 1.0)


def test_bpsk_llr_moments():
    """LLR = 2y/s2 with y = x + s n: times (1 - 2c) it is N(2/s2, 4/s2)."""
    print("\n[generator] BPSK moments", flush=True)
    code, ref = F.NrLdpc.lifted(2, 52), O.Code.lifted(2, 52)
    llr, info = F.generate_batch(code, 400, F.BPSK, 2.0, seed=3)
    llr = llr.cpu().numpy().astype(np.float64)
    bits = F.unpack_info(info, code.k)
    tx = np.array([ref.encode(b)[104:] for b in bits])
    v = llr * (1.0 - 2.0 * tx)
    s2 = O.lib.or_bpsk_noise_variance(2.0, 520 / 2600)
    assert abs(v.mean() - 2 / s2) < 0.01 * 2 / s2
    assert abs(v.var() - 4 / s2) < 0.02 * 4 / s2


def oracle_fer(code, ebn0, trials, iters=20):
    """The reference trial chain on the CPU: xoshiro streams, BPSK, f64 layered NMS 0.75."""
    fulls, cws = [], []
    for t in range(trials):
        cw, llr = code.trial_bpsk_lifted(ebn0, 1, "ref", 0, t)
        fulls.append(llr)
        cws.append(cw)
    post, its = code.decode_many(np.array(fulls), rule=O.NMS, alpha=0.75, max_iterations=iters, threads=8)
    errs = ((post[:, :code.k] < 0).astype(np.uint8) != np.array(cws)[:, :code.k]).sum(1)
    return int((errs > 0).sum()), errs, float(its.mean())


def test_campaign_fer_matches_the_reference_chain():
    print("\n[campaign] BG2 Z=52 BPSK 1.25 dB vs the oracle", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    res = F.Runner().seed(7).batch(4096).stop(F.Stop.max_trials(16384)).run_nr(
        code, F.Grid().axis("ebn0_db", [1.25]), F.BPSK, F.EBN0_DB)
    fer = res[0].metric("fer")
    ber = res[0].metric("ber")
    assert res[0].trials == 16384 and fer.total == 16384 and ber.total == 16384 * 520
    ref = O.Code.lifted(2, 52)
    fe, errs, it = oracle_fer(ref, 1.25, 3000)
    lo, hi = O.wilson(fe, 3000, 0.95)
    assert lo <= fer.value <= hi, (fer.value, fe / 3000, lo, hi)
    frac = errs / 520.0
    half = O.lib.or_z_score(0.95) * frac.std(ddof=1) / np.sqrt(frac.size)
    assert frac.mean() - half <= ber.value <= frac.mean() + half, (ber.value, frac.mean(), half)
    assert abs(res[0].metric("iterations").value - it) < 0.5


def test_campaign_records_do_not_depend_on_the_batch():
    print("\n[campaign] batch independence", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    grid = F.Grid().axis("ebn0_db", [1.0, 1.5])
    out = []
    for batch in (3000, 1000, 777):
        r = F.Runner().seed(5).batch(batch).stop(F.Stop.max_trials(3000)).run_nr(code, grid)
        out.append([(c.trials, c.metric("fer").count, c.metric("ber").count, c.metric("iterations").value)
                    for c in r])
    assert out[0] == out[1] == out[2]
    assert out[0][0][1] > out[0][1][1]  # more frame errors at the lower SNR


def test_campaign_stops_on_frame_errors():
    print("\n[campaign] min_errors stop", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 2).set_decoder("normalized_min_sum", 0.75, True, 10)
    r = F.Runner().seed(7).batch(512).stop(F.Stop.min_errors("fer", 50) | F.Stop.max_trials(20000)).run_nr(
        code, F.Grid().axis("ebn0_db", [0.5, 1.0]), F.QPSK, F.EBN0_DB)
    for c in r:
        m = c.metric("fer")
        assert c.status == "completed" and c.trials % 512 == 0 or c.trials == 20000
        assert m.count >= 50 or c.trials == 20000


def test_campaign_with_the_learned_decoder_and_its_fallback():
    """The learned link with the plain parameters (a = alpha, b = 0) and T = the code's limit
    decodes every frame as the plain link; the fallback decodes the frames that did not
    converge again, from their channel LLRs, with the same rule: the same decisions (equal
    bit and frame error counts) and T more iterations on each of those frames."""
    print("\n[campaign] learned link and fallback", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 12)
    grid = F.Grid().axis("ebn0_db", [1.0, 1.4])
    stop = F.Stop.max_trials(4096)
    plain = F.Runner().seed(9).batch(1024).stop(stop).run_nr(code, grid)
    p = F.initial_params(code, 12)
    same = F.Runner().seed(9).batch(1024).stop(stop).run_nr(code, grid, learned_params=p)
    fb = F.Runner().seed(9).batch(1024).stop(stop).run_nr(code, grid, learned_params=p, fallback=True)
    for a, b, c in zip(plain, same, fb):
        assert a.trials == b.trials == c.trials == 4096
        counts = [(r.metric("fer").count, r.metric("ber").count) for r in (a, b, c)]
        assert counts[0] == counts[1] == counts[2]
        assert b.metric("iterations").value == a.metric("iterations").value
        extra = (c.metric("iterations").value - a.metric("iterations").value) * 4096 / 12
        assert abs(extra - round(extra)) < 1e-6 * 4096  # whole frames of 12 more iterations
        assert round(extra) >= a.metric("fer").count > 0


def test_campaign_reports_and_skip_after(tmp_path):
    """The GPU campaign with the runner's observers: SkipAfter ends the SNR sweep once the
    FER is below 1e-3 (callbacks.rs:168-203), and the CSV rows and the JSON report
    (loggers.rs:107-318) hold the results the campaign returns."""
    import json
    print("\n[campaign] loggers and skip-after", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    grid = F.Grid().axis("ebn0_db", [1.0, 4.0, 5.0, 6.0])
    csv_path, json_path = tmp_path / "c.csv", tmp_path / "c.json"
    r = (F.Runner().seed(3).batch(2048).stop(F.Stop.max_trials(4096))
         .callback(F.SkipAfter.below("fer", 1e-3, "ebn0_db"))
         .logger(F.CsvLogger(csv_path)).logger(F.JsonLogger(json_path)).run_nr(code, grid))
    assert [c.status for c in r] == ["completed", "completed", "skipped", "skipped"]
    assert r[0].metric("fer").count > 0 and r[1].metric("fer").count == 0
    lines = csv_path.read_text().splitlines()
    assert lines[0] == ("index,ebn0_db,status,trials,wall_s," + ",".join(
        f"{m}{s}" for m in ("ber", "fer", "iterations") for s in ("", "_lo", "_hi", "_count", "_total")))
    assert len(lines) == 5 and lines[3].startswith("2,5,skipped,0,")
    for c, line in zip(r[:2], lines[1:3]):
        cells = line.split(",")
        assert float(cells[5]) == c.metric("ber").value and int(cells[8]) == c.metric("ber").count
        assert float(cells[10]) == c.metric("fer").value and int(cells[13]) == c.metric("fer").count
    doc = json.loads(json_path.read_text())
    assert doc["scenario"] == "nr_ldpc" and doc["seed"] == 3 and doc["batch"] == 2048
    assert [x["status"] for x in doc["conditions"]] == [c.status for c in r]
    assert doc["conditions"][0]["metrics"]["fer"]["errors"] == r[0].metric("fer").count
