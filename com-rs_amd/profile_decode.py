"""A short driver for profilers and timings: decode one batch of the BG1 Z=384 16-QAM
workload a few times with the plain or the learned decoder (FP32 or FP16 storage), or the
uneven workload (codewords of every NR base graph and lifting size 16..384, BPSK/QPSK,
Eb/N0 uniform in [0, 3] dB, iteration limits 5..25) through one batch plan, or the DVB-S2
short rate-1/2 code (a code given by its rows: the lane-per-codeword kernel), or copy a
buffer of known size (the yardstick of the HBM byte counters). Prints the mean time of a
step measured with HIP events.

    python com-rs_amd/profile_decode.py --decoder plain --batch 4096 --steps 5 --esn0 4.2
    python com-rs_amd/profile_decode.py --decoder plain --half
    python com-rs_amd/profile_decode.py --uneven 20000 --half --graph
    python com-rs_amd/profile_decode.py --copy-mib 1024 --steps 5
    python com-rs_amd/profile_decode.py --dvbs2 20000 --ebn0 0.9 --steps 2
    python com-rs_amd/profile_decode.py --learned-sweep 6,7,8,9,10 --esn0 4.2

Under rocprofv3, e.g.:
    rocprofv3 --kernel-trace --stats -d OUT -o plain -- python3 com-rs_amd/profile_decode.py
    rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 com-rs_amd/profile_decode.py
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--decoder", choices=["plain", "learned"], default="plain")
    p.add_argument("--batch", type=int, default=4096)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--esn0", type=float, default=4.2)
    p.add_argument("--iterations", type=int, default=8, help="learned decoder iterations")
    p.add_argument("--train-steps", type=int, default=200)
    p.add_argument("--copy-mib", type=int, default=0)
    p.add_argument("--half", action="store_true", help="FP16 message storage")
    p.add_argument("--uneven", type=int, default=0, help="codewords of the uneven workload")
    p.add_argument("--graph", action="store_true", help="uneven: replay a captured HIP graph")
    p.add_argument("--tune", action="store_true", help="uneven: autotune the plan's configuration first")
    p.add_argument("--tune-storage", action="store_true", help="uneven: let the autotuner change the storage too")
    p.add_argument("--per-code-width", action="store_true", help="uneven: each code at its own width")
    p.add_argument("--overlap", action="store_true", help="uneven: the launches of a run side by side")
    p.add_argument("--dvbs2", type=int, default=0,
                   help="codewords of the DVB-S2 short rate-1/2 code (layered sum-product, FP32, 50 it)")
    p.add_argument("--ebn0", type=float, default=0.9, help="dvbs2: Eb/N0 in dB")
    p.add_argument("--learned-sweep", default="",
                   help="comma-separated iteration counts T: the learned decoder trained for each, then its FER and "
                        "mean iterations (with and without the plain fallback) against the plain decoder and the "
                        "sum-product rule on the same frames")
    p.add_argument("--frames", type=int, default=32768, help="learned sweep: evaluation frames")
    p.add_argument("--sweep-train-steps", type=int, default=1000, help="learned sweep: Adam steps for each T")
    a = p.parse_args()
    import torch
    import comfec as F
    if a.learned_sweep:
        return learned_sweep(a, torch, F)
    if a.uneven:
        return uneven(a, torch, F)
    if a.dvbs2:
        return dvbs2(a, torch, F)
    if a.copy_mib:
        src = torch.empty(a.copy_mib << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(a.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {a.copy_mib} MiB x {a.steps}", flush=True)
        return
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    llr, info = F.generate_batch(code, a.batch, F.QAM16, a.esn0, snr_kind=F.ESN0_DB, seed=5)
    params = None
    if a.decoder == "learned":
        params = F.initial_params(code, a.iterations)
        F.learned_train(code, params, steps=a.train_steps, batch=512, lr=1e-2, modulation=F.QAM16,
                        snr_kind=F.ESN0_DB, snr_lo=a.esn0 - 0.1, snr_hi=a.esn0 + 0.3, seed=11)
    def step():
        if params is None:
            return code.decode_batch(llr, half=a.half)[1]
        return F.learned_decode_batch(code, params, llr, fallback=True)[1]

    its = step()
    ms = timed(step, a.steps, torch)
    print(f"{a.decoder}{' fp16' if a.half else ''}: {a.steps} x {a.batch} codewords, mean iterations "
          f"{its.float().mean().item():.3f}, {ms:.3f} ms per step, {8448 * a.batch / ms / 1e6:.3f} Gbit/s",
          flush=True)


def timed(fn, steps, torch):
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def learned_sweep(a, torch, F):
    """The learned decoder at several iteration counts T on the BG1 Z=384 16-QAM workload at
    Es/N0 `esn0`: trained on the device as bench.py trains it (1000 Adam steps of 512
    codewords, lr 1e-2, Es/N0 uniform in [esn0 - 0.1, esn0 + 0.3]), then decoded on `frames`
    fresh frames with and without the plain fallback; the plain NMS 0.75 with 10 iterations,
    and the layered NMS and FP32 sum-product with a limit of 50 on the same frames. One JSON
    line per decoder: frames, frame errors, FER with its 95 % Wilson interval, mean
    iterations, and the time of one 4096-codeword launch (HIP events)."""
    import json
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    nms50 = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 50)
    spa50 = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("sum_product", 0.0, True, 50)
    parts = [F.generate_batch(code, 4096, F.QAM16, a.esn0, snr_kind=F.ESN0_DB, seed=6, point=0, trial0=4096 * i)
             for i in range(a.frames // 4096)]

    def evaluate(name, fn):
        fe, its_sum = 0, 0.0
        for llr, info in parts:
            bits, its = fn(llr)
            fe += int((bits.cpu().numpy() != F.unpack_info(info, 8448)).any(1).sum())
            its_sum += float(its.double().sum().item())
        n = len(parts) * 4096
        ms = timed(lambda: fn(parts[0][0]), 5, torch)
        row = {"decoder": name, "esn0_db": a.esn0, "frames": n, "frame_errors": fe, "fer": fe / n,
               "fer_wilson95": [round(x, 8) for x in F.wilson(fe, n, 0.95)], "mean_iterations": round(its_sum / n, 4),
               "ms_per_4096": round(ms, 4), "gbit_s": round(8448 * 4096 / ms / 1e6, 3)}
        print(json.dumps(row), flush=True)
        return row

    evaluate("plain NMS 0.75, 10 it", lambda l: code.decode_batch(l)[:2])
    evaluate("NMS 0.75, limit 50", lambda l: nms50.decode_batch(l)[:2])
    evaluate("sum-product FP32, limit 50", lambda l: spa50.decode_batch(l)[:2])
    for T in [int(t) for t in a.learned_sweep.split(",")]:
        params = F.initial_params(code, T)
        F.learned_train(code, params, steps=a.sweep_train_steps, batch=512, lr=1e-2,
                        modulation=F.QAM16, snr_kind=F.ESN0_DB, snr_lo=a.esn0 - 0.1, snr_hi=a.esn0 + 0.3, seed=11)
        evaluate(f"learned T={T}", lambda l: F.learned_decode_batch(code, params, l, fallback=False)[:2])
        evaluate(f"learned T={T} + plain fallback", lambda l: F.learned_decode_batch(code, params, l, fallback=True)[:2])


LIFTS = [16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128,
         144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384]


def dvbs2(a, torch, F):
    """The DVB-S2 LDPC code of the short FECFRAME, rate 1/2 (16200 bits, 9000 checks), built
    from the standard's address table (tests/golden/dvbs2_n16200_k7200.json) by the rule of
    crates/com-systems/src/dvbs2/fec.rs:213-243; BPSK over AWGN, the all-zero codeword, LLRs
    from torch's generator; one batch decoded `steps` times by the lane-per-codeword kernel."""
    import json
    import numpy as np
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    fx = json.load(open(os.path.join(root, "tests", "golden", "dvbs2_n16200_k7200.json")))
    n, k, group, table = fx["n"], fx["k"], fx["group"], fx["table"]
    m, q = n - k, (n - k) // group
    rows = [[] for _ in range(m)]
    bit = i = 0
    while i < len(table):
        for l in range(group):
            for x in table[i + 1:i + 1 + table[i]]:
                rows[(x + l * q) % m].append(bit)
            bit += 1
        i += 1 + table[i]
    for r in range(m):
        rows[r] += ([k + r - 1] if r > 0 else []) + [k + r]
    code = F.NrLdpc.from_rows(n, rows).set_decoder("sum_product", 0.0, True, 50)
    s2 = 1.0 / (2.0 * (k / n) * 10.0 ** (a.ebn0 / 10.0))
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    noise = torch.randn((a.dvbs2, n), generator=g, dtype=torch.float32, device="cuda")
    llr = ((1.0 + float(np.sqrt(s2)) * noise) * float(2.0 / s2)).contiguous()
    del noise
    code.decode_batch(llr)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(a.steps):
        bits, its, _ = code.decode_batch(llr)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / a.steps
    it = float(its.double().mean().item())
    print(f"dvbs2 short 1/2: {a.dvbs2} codewords at {a.ebn0} dB, mean iterations {it:.3f}, "
          f"max {int(its.max().item())}, {ms:.3f} ms per launch, {a.dvbs2 / ms * 1e3:.0f} codewords/s", flush=True)


def uneven(a, torch, F):
    """The uneven workload: descriptors from a fixed seed, LLRs from the device generator."""
    import numpy as np
    rng = np.random.default_rng(7)
    n = a.uneven
    bgs = rng.integers(1, 3, n)
    zs = np.array(LIFTS)[rng.integers(0, len(LIFTS), n)]
    mods = np.where(rng.integers(0, 2, n) == 0, F.BPSK, F.QPSK)
    limits = rng.integers(5, 26, n).astype(np.uint32)
    codes = {}
    for b, z in set(zip(bgs.tolist(), zs.tolist())):
        codes[(b, z)] = F.NrLdpc.lifted(b, z).set_decoder("normalized_min_sum", 0.75, True, 25)
    groups = {}
    for i in range(n):
        groups.setdefault((int(bgs[i]), int(zs[i]), int(mods[i])), []).append(i)
    per = [None] * n
    for gi, ((b, z, m), idx) in enumerate(sorted(groups.items())):
        llr, _ = F.generate_training_batch(codes[(b, z)], len(idx), m, 0.0, 3.0, snr_kind=F.EBN0_DB, seed=7,
                                           point=gi)
        for j, i in enumerate(idx):
            per[i] = llr[j]
    cl = [codes[(int(b), int(z))] for b, z in zip(bgs, zs)]
    offsets = np.zeros(n + 1, np.uint64)
    offsets[1:] = np.cumsum([c.dims.e for c in cl])
    flat = torch.cat(per).contiguous()
    plan = F.BatchPlan(cl, offsets, limits, half=a.half)
    out = plan.outputs()
    if a.per_code_width:
        plan.configure(per_code_width=True, overlap=a.overlap)
    if a.tune:
        best, _ = plan.tune(flat, out, repeats=3, storage=a.tune_storage)
        for t in plan.trials:
            print(f"  trial {t.config!r}: {t.segments} launch(es), {t.ms:.3f} ms", flush=True)
        print(f"tuned: {plan.config!r} ({best:.3f} ms)", flush=True)
    plan.run(flat, out=out, graph=a.graph)
    ms = timed(lambda: plan.run(flat, out=out, graph=a.graph), a.steps, torch)
    d = plan.dims
    its = out[1].double()
    launches = ", ".join(f"{sg.grid} x {sg.threads} lanes ({sg.classes} codes, {sg.tasks} tasks, LDS {sg.lds_bytes} B)"
                         for sg in plan.segments())
    print(f"uneven{' fp16' if d.msg_type == F.MSG_F16 else ''}{' graph' if a.graph else ''}: {n} codewords, "
          f"{d.classes} codes, {d.tasks} tasks, launches: {launches}; mean iterations "
          f"{its.mean().item():.3f}, {ms:.3f} ms per run, {n / ms * 1e3:.0f} codewords/s, "
          f"{d.bits_total / ms / 1e6:.3f} Gbit/s", flush=True)


if __name__ == "__main__":
    main()
