"""Throughput of the NR LDPC decoders on one MI355X at the FER 1e-3 operating point:
the learned offset/scale min-sum against the plain normalized min-sum.

Workload: NR base graph 1, Z=384 (K'=8448, E=25344, rate 1/3), Gray 16-QAM over AWGN with
max-log LLRs and the NR bit interleaver, 4096 codewords per step, FP32 messages, layered
schedule with syndrome early stop. The plain decoder is the normalized min-sum (alpha 0.75)
with 10 iterations. The operating point is the Es/N0 at which the plain decoder's FER is
nearest 1e-3 (measured on a grid of 0.1 dB over 16384 frames per point). The learned
decoder (per-iteration, per-base-row offset and scale, at most LEARNED_ITERATIONS
iterations, codewords whose syndrome still fails decoded again by the plain decoder) is
trained on the GPU before the timed steps, on LLRs of the device generator around that
point. A step decodes one batch whose LLRs are already resident in HBM; the timed steps
decode the same batch with each decoder.

Prints one JSON line; `value` is the learned decoder's decoded Gbit/s.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "com-rs_amd"))

# MI355X published peaks: HBM3E 8 TB/s; LDS 128 B/clk per CU x 256 CUs x 2.4 GHz; VALU one
# wave64 instruction per 4 clocks per SIMD x 1024 SIMDs x 2.4 GHz
HBM_PUBLISHED_GBS = 8000.0
LDS_PUBLISHED_GBS = 256 * 128 * 2.4
VALU_PUBLISHED_GINST = 1024 * 2.4 / 4
# What the profiles of the plain launch on this workload measured, per codeword
# (profiles/decode_layered_bg1_z384_16qam.md): HBM bytes from rocprofv3 FETCH_SIZE (x 2.0)
# and WRITE_SIZE (x 1.0), the channel LLRs in and the decisions out (453.4 MB per 4096
# codewords, 1.6 MB of it the stores of two registers spilled in the task set-up);
# SQ_INSTS_VALU per codeword and iteration (both wave sets, 12 waves)
TRAFFIC_PER_CODEWORD = 110690
VALU_PER_CODEWORD_ITERATION = 32144
PROFILE = "profiles/decode_layered_bg1_z384_16qam.md"
LEARNED_ITERATIONS = 10
K_INFO, E, Z, BG1_EDGES = 8448, 25344, 384, 316


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--train-steps", type=int, default=1000)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    return p.parse_args()


def measure_peaks(torch):
    """The HBM, LDS and VALU peaks of this device, by the microbenchmarks of
    com-rs_amd/csrc/peaks.hip (mean of 10 launches after a warm-up each)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "com-rs_amd", "_build", "libcomfec_peaks.so"))
    for name in ("cf_peak_hbm_read", "cf_peak_hbm_copy", "cf_peak_lds", "cf_peak_valu"):
        getattr(lib, name).restype = C.c_double
    sink = torch.zeros(1 << 16, dtype=torch.int32, device="cuda")
    src = torch.ones(1 << 29, dtype=torch.int32, device="cuda")  # 2 GiB
    dst = torch.empty(1 << 28, dtype=torch.int32, device="cuda")  # 1 GiB
    w = C.c_double()
    sp = C.c_void_p(sink.data_ptr())
    out = {}
    ms = lib.cf_peak_hbm_read(C.c_void_p(src.data_ptr()), C.c_size_t(src.numel() * 4), sp, 10, C.byref(w))
    out["hbm_read_gbs"] = round(w.value / ms / 1e6, 1)
    ms = lib.cf_peak_hbm_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), C.c_size_t(dst.numel() * 4),
                              10, C.byref(w))
    out["hbm_copy_gbs"] = round(w.value / ms / 1e6, 1)
    del src, dst
    for width in (4, 16):
        ms = lib.cf_peak_lds(width, 4096, sp, 10, C.byref(w))
        out[f"lds_read_b{8 * width}_gbs"] = round(w.value / ms / 1e6, 1)
    for kind, name in ((0, "valu_ginst_s"), (1, "valu_int_add_ginst_s"), (2, "valu_bitop3_ginst_s")):
        ms = lib.cf_peak_valu(kind, 20000, sp, 10, C.byref(w))
        out[name] = round(w.value / ms / 1e6, 1)
    torch.cuda.synchronize()
    return out


def host_threads():
    """The host threads the CPU baseline may use: the CPUs this process may run on, at most
    the box's share given by OMP_NUM_THREADS when that is set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(llr_f32, seconds, esn0):
    """The oracle's f64 layered NMS (a C restatement of LdpcCode::propagate, 10 iterations,
    syndrome stop) on host threads over a bounded sample of the timed batch, on all cores
    and on one; and, once, the one published timing of the reference's decoder
    (benchmarks/libraries/results/s4_coding.json:248,333: NR K'=1024 E=2048, BG2 Z=104,
    flooding sum-product f64, at most 50 iterations, 200 frames at Eb/N0 2.0 dB, 0.34056 s
    on an Apple M6 thread) for the ratio."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ref = O.Code.nr(K_INFO, E, 1, 0, 4)
    threads = host_threads()

    def rate(count, nthreads):
        fulls = np.array([ref.full_llr(l.astype(np.float64)) for l in llr_f32[:count]])
        t0 = time.perf_counter()
        ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=nthreads)
        return time.perf_counter() - t0

    probe = min(threads, llr_f32.shape[0])
    dt = rate(probe, threads)
    count = int(max(probe, min(llr_f32.shape[0], probe * seconds / max(dt, 1e-3))))
    dt = rate(count, threads)
    one = max(1, min(count, int(2.0 * count / max(dt * threads, 1e-3))))  # about 2 s on one thread
    dt1 = rate(one, 1)

    # the published workload, same chain as the reference's benchmark: random bits, the NR
    # encoder, BPSK over AWGN at the rate K'/E, f64 LLRs 2y/sigma^2
    cal = O.Code.nr(1024, 2048)
    s2 = O.lib.or_bpsk_noise_variance(2.0, 1024 / 2048)
    fulls = []
    for t in range(200):
        r = O.Rng.stream(1, "s4_coding", 0, t)
        info = np.array([r.bit() for _ in range(1024)], np.uint8)
        llr = np.zeros(2048)
        O.lib.or_bpsk_awgn_llr(cal.nr_encode(info), 2048, s2, r.r, llr)
        fulls.append(cal.full_llr(llr))
    fulls = np.array(fulls)
    cal.decode_many(fulls, rule=O.SPA, alpha=0.0, layered=False, max_iterations=50, threads=1)
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        cal.decode_many(fulls, rule=O.SPA, alpha=0.0, layered=False, max_iterations=50, threads=1)
        runs.append(time.perf_counter() - t0)
    cal_s = float(np.median(runs))

    # the whole trial chain on the same threads: xoshiro streams, bits, NR encoder, rate
    # matching and interleaver, Gray 16-QAM, AWGN, max-log LLRs, rate recovery, the decoder,
    # the bit and frame error counts
    n0 = 1.0 / O.lib.or_db2lin(esn0)
    chain = max(threads, count // 2)
    t0 = time.perf_counter()
    info, llr = ref.trial_nr_qam_many(16, n0, 1, "bench", 0, 0, chain, threads=threads)
    fulls = np.array([ref.full_llr(l) for l in llr])
    post, _ = ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=threads)
    errs = ((post[:, :K_INFO] < 0).astype(np.uint8) != info).sum(1)
    dt_chain = time.perf_counter() - t0
    return {"value": round(K_INFO * count / dt / 1e9, 6), "unit": "Gbit/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "full_chain": {"value": round(K_INFO * chain / dt_chain / 1e9, 6), "unit": "Gbit/s",
                           "sample": f"{chain} trials at Es/N0 {esn0} dB (bits, encoder, interleaver, 16-QAM, "
                                     f"AWGN, max-log LLRs, decoder, error counts) on {threads} threads, "
                                     f"{dt_chain:.1f} s, {int((errs > 0).sum())} frame errors"},
            "sample": f"{count} codewords of the timed batch: f64 layered NMS 0.75, 10 it, syndrome stop "
                      f"(oracle C restatement of LdpcCode::propagate), {threads} threads, {dt:.1f} s",
            "single_thread": {"value": round(K_INFO * one / dt1 / 1e9, 6), "unit": "Gbit/s",
                              "sample": f"{one} codewords, 1 thread, {dt1:.2f} s"},
            "calibration": {"workload": "NR K'=1024 E=2048 (BG2 Z=104), flooding sum-product f64, <= 50 it, "
                                        "200 frames at Eb/N0 2.0 dB, 1 thread, median of 5",
                            "seconds": round(cal_s, 4), "reference_seconds_apple_m6": 0.34056,
                            "ratio_to_reference": round(cal_s / 0.34056, 3)}}


def frame_errors(bits, info, F):
    import numpy as np
    return int((bits.cpu().numpy() != F.unpack_info(info, bits.shape[1])).any(1).sum())


def timed(fn, steps, warmup, torch):
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps, time.perf_counter() - t0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the decoder shards by "
              "codeword; see the multi-GPU paragraph of the design notes)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import comfec as F

    torch.cuda.init()
    code = F.NrLdpc(K_INFO, E, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    batch = args.batch

    # the operating point: plain FER nearest 1e-3 on a 0.1 dB grid (16384 frames per point)
    calib = {}
    for i in range(8):
        esn0 = round(3.9 + 0.1 * i, 2)
        fe, n = 0, 0
        for part in range(4):
            llr, info = F.generate_batch(code, 4096, F.QAM16, esn0, snr_kind=F.ESN0_DB, seed=2, point=i,
                                         trial0=4096 * part)
            bits, _, _ = code.decode_batch(llr)
            fe += frame_errors(bits, info, F)
            n += 4096
        calib[esn0] = (fe, n)
    esn0 = min(calib, key=lambda p: abs(math.log10(max(calib[p][0], 0.5) / calib[p][1]) + 3.0))
    # the config grid of the workload, for the record
    grid = {}
    for i, p in enumerate((4.0, 5.0, 6.0, 7.0)):
        llr, info = F.generate_batch(code, batch, F.QAM16, p, snr_kind=F.ESN0_DB, seed=3, point=i)
        bits, its, _ = code.decode_batch(llr)
        grid[str(p)] = {"fer": frame_errors(bits, info, F) / batch, "mean_iterations":
                        round(float(its.float().mean().item()), 3)}

    # the learned decoder, trained on the device around the operating point
    params = F.initial_params(code, LEARNED_ITERATIONS)
    t0 = time.perf_counter()
    hist = F.learned_train(code, params, steps=args.train_steps, batch=512, lr=1e-2, modulation=F.QAM16,
                           snr_kind=F.ESN0_DB, snr_lo=esn0 - 0.1, snr_hi=esn0 + 0.3, seed=11)
    train_s = time.perf_counter() - t0

    # the timed batch and an independent evaluation set at the operating point
    llr, info = F.generate_batch(code, batch, F.QAM16, esn0, snr_kind=F.ESN0_DB, seed=5, point=0)
    ev = {"plain": [0, 0, 0.0], "learned": [0, 0, 0.0]}
    for part in range(8):
        l2, i2 = F.generate_batch(code, 4096, F.QAM16, esn0, snr_kind=F.ESN0_DB, seed=6, point=0, trial0=4096 * part)
        b, its, _ = code.decode_batch(l2)
        ev["plain"][0] += frame_errors(b, i2, F)
        ev["plain"][1] += 4096
        ev["plain"][2] += float(its.double().sum().item())
        b, its, _ = F.learned_decode_batch(code, params, l2, fallback=True)
        ev["learned"][0] += frame_errors(b, i2, F)
        ev["learned"][1] += 4096
        ev["learned"][2] += float(its.double().sum().item())

    _, its_plain, _ = code.decode_batch(llr)
    _, its_learned, _ = F.learned_decode_batch(code, params, llr, fallback=True)
    it_plain = float(its_plain.float().mean().item())
    it_learned = float(its_learned.float().mean().item())
    ms_plain, wall_plain = timed(lambda: code.decode_batch(llr), args.steps, args.warmup, torch)
    ms_learned, wall_learned = timed(lambda: F.learned_decode_batch(code, params, llr, fallback=True),
                                     args.steps, args.warmup, torch)
    gbps_plain = K_INFO * batch / (ms_plain / 1e3) / 1e9
    gbps_learned = K_INFO * batch / (ms_learned / 1e3) / 1e9

    # roofline of the layered kernel (one launch per plain step), against the resource that
    # binds it: the posteriors live in LDS and the messages in registers, so its algorithmic
    # bytes are LDS bytes, I x E_dg x 16 per codeword (posterior read and write, variable-to-
    # check message write and read, 4 B each); the unit is one codeword, a launch decodes
    # `batch` of them, at the mean iteration count I of the timed batch
    peaks = measure_peaks(torch)
    edges = BG1_EDGES * Z
    s_plain = ms_plain / 1e3
    lds_bytes = batch * it_plain * edges * 16
    lds_achieved = lds_bytes / s_plain / 1e9
    lds_peak = peaks["lds_read_b32_gbs"]
    valu_rate = batch * it_plain * VALU_PER_CODEWORD_ITERATION / s_plain / 1e9
    hbm_yard = batch * (E * 4 + math.ceil(K_INFO / 8) + it_plain * 2 * edges * 4)
    wil = lambda fe, n: [round(x, 7) for x in F.wilson(fe, n, 0.95)]
    out = {
        "metric": "decoded Gbit/s at FER 1e-3, learned vs plain min-sum, NR BG1 Z=384",
        "value": round(gbps_learned, 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_learned, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: device Philox trial generator (random info bits, NR encoder + rate matching + "
                "interleaver, Gray 16-QAM, AWGN, max-log LLRs)",
        "config": {
            "workload": "NR BG1 Z=384 K=8448 E=25344 rate 1/3, 16-QAM, layered min-sum, 4096 codewords/step, "
                        "Es/N0 at plain FER 1e-3",
            "global_batch": batch,
            "esn0_db": esn0,
            "plain": {"decoder": "normalized min-sum a=0.75, 10 it", "gbit_s": round(gbps_plain, 4),
                      "ms_per_step": round(ms_plain, 4), "mean_iterations": round(it_plain, 3),
                      "fer": ev["plain"][0] / ev["plain"][1], "fer_wilson95": wil(*ev["plain"][:2]),
                      "frames": ev["plain"][1]},
            "learned": {"decoder": f"offset/scale min-sum, {LEARNED_ITERATIONS} it, plain fallback",
                        "gbit_s": round(gbps_learned, 4), "ms_per_step": round(ms_learned, 4),
                        "mean_iterations": round(it_learned, 3),
                        "fer": ev["learned"][0] / ev["learned"][1], "fer_wilson95": wil(*ev["learned"][:2]),
                        "frames": ev["learned"][1], "train_steps": args.train_steps,
                        "train_s": round(train_s, 2), "loss_first_last": [round(float(hist[:20].mean()), 5),
                                                                          round(float(hist[-20:].mean()), 5)]},
            "speedup_learned_vs_plain": round(gbps_learned / gbps_plain, 4),
            "iteration_ratio_learned_vs_plain": round(it_learned / it_plain, 4),
            "plain_fer_calibration": {str(p): [fe, n] for p, (fe, n) in calib.items()},
            "plain_fer_by_esn0_grid": grid,
            "wall_s_timed": round(wall_plain + wall_learned, 4),
        },
        "roofline": {"bound": "lds", "achieved": round(lds_achieved, 1), "peak": lds_peak, "unit": "GB/s",
                     "frac": round(lds_achieved / lds_peak, 4), "traffic": TRAFFIC_PER_CODEWORD * batch,
                     "kernel": "decode_layered_dual_kernel<1, true> (plain step)",
                     "unit_of_work": f"one codeword, I x E_dg x 16 B of LDS (I = {it_plain:.3f}, E_dg = {edges}); "
                                     f"{batch} a launch",
                     "peak_source": "measured: conflict-free 4-byte LDS reads on every CU (peaks.hip)",
                     "peaks_measured": peaks,
                     "peaks_published": {"hbm_gbs": HBM_PUBLISHED_GBS, "lds_gbs": LDS_PUBLISHED_GBS,
                                         "valu_ginst_s": VALU_PUBLISHED_GINST},
                     "valu": {"ginst_s": round(valu_rate, 1), "peak": peaks["valu_ginst_s"],
                              "frac": round(valu_rate / peaks["valu_ginst_s"], 4),
                              "note": f"wave instructions per codeword and iteration from SQ_INSTS_VALU ({PROFILE}) "
                                      "times the live iterations and batch over the live step time: the issue "
                                      "rate that bounds the kernel"},
                     "traffic_source": f"{PROFILE}: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (a 1 GiB copy's "
                                       "calibration), the LLRs in and the decisions out",
                     "hbm_yardstick": {"bytes": round(hbm_yard), "achieved_gbs": round(hbm_yard / s_plain / 1e9, 1),
                                       "note": "the streaming decoder's HBM bytes (LLRs + K'/8 + I x 2 x E_dg x 4 per "
                                               "codeword); this kernel keeps those messages on chip, so the figure "
                                               "exceeds the HBM peak and is no roofline"}},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(llr.cpu().numpy(), args.cpu_seconds, esn0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
