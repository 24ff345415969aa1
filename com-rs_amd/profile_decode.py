"""A short driver for profilers: decode one batch of the BG1 Z=384 16-QAM workload a few
times with the plain or the learned decoder (or copy a buffer of known size, the yardstick
of the HBM byte counters).

    python com-rs_amd/profile_decode.py --decoder plain --batch 4096 --steps 5 --esn0 4.2
    python com-rs_amd/profile_decode.py --copy-mib 1024 --steps 5

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
    a = p.parse_args()
    import torch
    import comfec as F
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
    its = None
    for _ in range(a.steps):
        if params is None:
            _, its, _ = code.decode_batch(llr)
        else:
            _, its, _ = F.learned_decode_batch(code, params, llr, fallback=True)
    torch.cuda.synchronize()
    print(f"{a.decoder}: {a.steps} x {a.batch} codewords, mean iterations {its.float().mean().item():.3f}",
          flush=True)


if __name__ == "__main__":
    main()
