"""Synthetic inputs of the NR BG1 Z=384 decoder workload, built on the GPU.

Random information bits, the NR encoder and rate matching of the library
(cf_encode_block: BlockCode::encode_block of nr/ldpc.rs:208-214), Gray square QAM of unit
energy (constellation.rs:168-205, 396-398), complex AWGN of variance N0 per symbol
(awgn.rs:94-99) and max-log LLRs with the uniform prior term (demapper.rs:270-297).
The channel runs in torch on the device (float64, then stored as float32), so that the
LLR batch is resident in HBM before any timed decode.
"""
import numpy as np

import comfec as F


def qam_points(m):
    """Gray-labelled square QAM, points[label], unit average energy."""
    l = int(round(np.sqrt(m)))
    lv = 2.0 * np.arange(l) - (l - 1)
    nat = []
    for r in range(l):
        row = [complex(re, lv[l - 1 - r]) for re in lv]
        nat.extend(row[::-1] if r % 2 else row)
    pts = np.zeros(m, complex)
    for i, p in enumerate(nat):
        pts[i ^ (i >> 1)] = p
    return pts / np.sqrt(np.mean(np.abs(pts) ** 2))


def encode_batch(code, count, seed):
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (count, code.k), dtype=np.uint8)
    tx = np.stack([code.encode_block(u) for u in info])
    return info, tx


def qam_llr_torch(tx_bits, m, n0, generator, chunk=512):
    """Map (count, E) bits to QAM, add noise of variance n0, max-log LLRs (count, E) float32."""
    import torch
    dev = tx_bits.device
    b = int(np.log2(m))
    pts = torch.from_numpy(qam_points(m)).to(dev)
    out = torch.empty(tx_bits.shape, dtype=torch.float32, device=dev)
    weights = (2 ** torch.arange(b - 1, -1, -1, device=dev)).to(torch.int64)
    labels = torch.arange(m, device=dev)
    masks = [((labels >> (b - 1 - o)) & 1).bool() for o in range(b)]
    uniform = -np.log(m)
    for s in range(0, tx_bits.shape[0], chunk):
        bits = tx_bits[s:s + chunk].to(torch.int64).reshape(-1, tx_bits.shape[1] // b, b)
        idx = (bits * weights).sum(-1)
        x = pts[idx]
        noise = torch.complex(torch.randn(x.shape, generator=generator, device=dev, dtype=torch.float64),
                              torch.randn(x.shape, generator=generator, device=dev, dtype=torch.float64))
        y = x + noise * np.sqrt(n0 / 2.0)
        metric = -(y.unsqueeze(-1) - pts).abs().square() / n0 + uniform
        llr = torch.empty(bits.shape, dtype=torch.float64, device=dev)
        for o in range(b):
            m0 = metric[..., ~masks[o]].amax(-1)
            m1 = metric[..., masks[o]].amax(-1)
            llr[..., o] = m0 - m1
        out[s:s + chunk] = llr.reshape(bits.shape[0], -1).to(torch.float32)
    return out


def nr_qam_batch(code, count, esn0_db, m=16, seed=1, device="cuda"):
    """(info bits (count, K') numpy, LLRs (count, E) float32 on the device) at Es/N0."""
    import torch
    info, tx = encode_batch(code, count, seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed * 1000 + int(esn0_db * 10))
    n0 = 10.0 ** (-esn0_db / 10.0)
    llr = qam_llr_torch(torch.from_numpy(tx).to(device), m, n0, gen)
    return info, llr
