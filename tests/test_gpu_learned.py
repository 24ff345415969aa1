"""The learned offset/scale min-sum on the GPU (no reference counterpart), through the C ABI.

- with the plain parameters it decodes exactly as the plain layered min-sum kernel;
- its loss and gradient (hand-written forward and backward kernels, FP32) match the f64
  oracle, whose backward is pinned by finite differences (tests/test_oracle_learned.py);
- training by the device generator, the kernels and Adam lowers the loss, and the trained
  decoder has a lower frame error rate than the plain normalized min-sum at the same
  iteration count; the fallback hands the frames whose syndrome failed to the plain rule.
"""
import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_plain_parameters_decode_as_the_plain_kernel():
    print("\n[learned] plain parameters = plain kernel", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 12)
    llr, _ = F.generate_batch(code, 256, F.BPSK, 1.0, seed=4)
    bits, its, _ = code.decode_batch(llr)
    p = F.initial_params(code, 12)
    bits2, its2, ok = F.learned_decode_batch(code, p, llr, fallback=False)
    assert torch.equal(bits, bits2) and torch.equal(its, its2)
    # offset min-sum too
    code.set_decoder("offset_min_sum", 0.15, True, 12)
    bits, its, _ = code.decode_batch(llr)
    bits2, its2, _ = F.learned_decode_batch(code, F.initial_params(code, 12, 1.0, 0.15), llr, fallback=False)
    assert torch.equal(bits, bits2) and torch.equal(its, its2)


def test_bg1_z384_plain_parameters_and_fallback_decode_as_the_plain_kernel():
    """BG1 Z=384 (the two-wave-set kernel with Z compiled in, whose iteration loop takes
    the codeword's state from its first lane): the learned launch with the plain parameters
    (a = 0.75, b = 0) equals the plain normalized min-sum launch bit for bit; with T = 3 and
    the fallback, a frame that the three learned iterations do not finish is decoded again
    from its LLRs by the plain rule, so every frame's decisions are the plain decoder's, and
    its iteration count is the plain count (at most 3) or 3 plus it."""
    print("\n[learned] BG1 Z=384 16-QAM: plain parameters, T = 3 with fallback", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    llr, _ = F.generate_batch(code, 512, F.QAM16, 4.2, snr_kind=F.ESN0_DB, seed=21)
    bits, its, ok = code.decode_batch(llr)
    bits2, its2, ok2 = F.learned_decode_batch(code, F.initial_params(code, 10), llr, fallback=False)
    assert torch.equal(bits, bits2) and torch.equal(its, its2)
    T = 3
    bits3, its3, ok3 = F.learned_decode_batch(code, F.initial_params(code, T), llr, fallback=True)
    assert torch.equal(bits3, bits)
    p = its.cpu().numpy()
    want = np.where(p <= T, p, T + p)
    assert np.array_equal(its3.cpu().numpy(), want)
    assert np.count_nonzero(p > T) > 100  # the fallback path runs on most frames
    # ok: finished within the learned iterations
    assert np.array_equal(ok3.cpu().numpy().astype(bool), p <= T)


@pytest.mark.parametrize("bg,z,T", [(2, 52, 4), (1, 16, 3)])
def test_loss_and_gradient_match_the_oracle(bg, z, T):
    print(f"\n[learned] gradient vs oracle BG{bg} Z={z} T={T}", flush=True)
    code = F.NrLdpc.lifted(bg, z)
    ref = O.Code.lifted(bg, z)
    rows = 46 if bg == 1 else 42
    count = 8
    llr, cw = F.generate_training_batch(code, count, F.BPSK, 0.5, 2.5, seed=9)
    g = torch.Generator().manual_seed(bg * 1000 + z)
    params = torch.stack([0.75 + 0.1 * torch.randn(T, rows, generator=g),
                          0.05 + 0.1 * torch.rand(T, rows, generator=g)], -1).float().cuda()
    loss, grad = F.learned_loss_grad(code, params, llr, cw)
    loss, grad = float(loss.item()), grad.cpu().numpy().astype(np.float64)
    bits = F.unpack_info(cw, ref.n)
    a = params[..., 0].cpu().numpy().astype(np.float64)
    b = params[..., 1].cpu().numpy().astype(np.float64)
    lref, ga, gb = 0.0, np.zeros((T, rows)), np.zeros((T, rows))
    llr_np = llr.cpu().numpy().astype(np.float64)
    for i in range(count):
        full = np.zeros(ref.n)
        full[2 * z:] = llr_np[i]
        l, ga_i, gb_i = ref.learned_loss_grad(full, bits[i], a, b)
        lref += l / count
        ga += ga_i / count
        gb += gb_i / count
    assert loss == pytest.approx(lref, rel=1e-4)
    ref_grad = np.stack([ga, gb], -1)
    scale = np.abs(ref_grad).max()
    assert scale > 0
    assert np.max(np.abs(grad - ref_grad)) <= 2e-3 * scale


def fer(bits, info):
    return float((bits.cpu().numpy() != F.unpack_info(info, bits.shape[1])).any(1).mean())


def test_training_lowers_the_loss_and_beats_the_plain_rule():
    print("\n[learned] training BG2 Z=52, 12 iterations", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 12)
    T = 12
    params = F.initial_params(code, T)
    hist = F.learned_train(code, params, steps=400, batch=4096, lr=1e-2, modulation=F.BPSK,
                           snr_lo=0.0, snr_hi=3.0, seed=1)
    assert np.all(np.isfinite(hist))
    print(f"training loss {hist[:50].mean():.5f} -> {hist[-50:].mean():.5f}", flush=True)
    assert hist[-50:].mean() < hist[:50].mean()
    # evaluation on fresh trials at 1.5 dB: learned vs plain, the same 12 iterations
    n = 40000
    llr, info = F.generate_batch(code, n, F.BPSK, 1.5, seed=12345)
    plain_bits, plain_its, _ = code.decode_batch(llr)
    bits, its, ok = F.learned_decode_batch(code, params, llr, fallback=False)
    f_plain, f_learned = fer(plain_bits, info), fer(bits, info)
    print(f"FER plain {f_plain:.4g} learned {f_learned:.4g}; iterations "
          f"{plain_its.float().mean().item():.3f} vs {its.float().mean().item():.3f}", flush=True)
    lo_p, hi_p = O.wilson(int(round(f_plain * n)), n, 0.95)
    lo_l, hi_l = O.wilson(int(round(f_learned * n)), n, 0.95)
    assert hi_l < lo_p
    # the fallback decodes again, with the plain rule, the frames whose syndrome failed
    bits_fb, its_fb, ok_fb = F.learned_decode_batch(code, params, llr, fallback=True)
    assert torch.equal(ok, ok_fb)
    okn = ok.cpu().numpy().astype(bool)
    assert torch.equal(bits_fb[torch.from_numpy(okn).cuda()], bits[torch.from_numpy(okn).cuda()])
    failed = ~okn
    assert np.all(its_fb.cpu().numpy()[failed] >= T)
    # the frames the fallback decoded are the plain decoder's own decisions
    pb = plain_bits.cpu().numpy()[failed]
    assert np.array_equal(bits_fb.cpu().numpy()[failed], pb)


def test_training_again_in_the_same_process_is_finite_and_repeatable():
    """Several trainings in one process with different iteration counts (8, 10, 7, then 8
    again), the sequence that once left Adam's moments uninitialised: every loss finite, and
    the second run at T = 8 repeats the first bit for bit (fixed seeds, fixed-order
    reductions)."""
    print("\n[learned] training T = 8, 10, 7, 8 in one process (BG1 Z=384 16-QAM)", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    runs = []
    for T in (8, 10, 7, 8):
        params = F.initial_params(code, T)
        hist = F.learned_train(code, params, steps=30, batch=256, lr=1e-2, modulation=F.QAM16, snr_kind=F.ESN0_DB,
                               snr_lo=4.1, snr_hi=4.5, seed=11)
        assert np.all(np.isfinite(hist)), (T, hist)
        assert torch.isfinite(params).all()
        runs.append((hist, params.cpu().numpy()))
    assert np.array_equal(runs[0][0], runs[3][0])
    assert np.array_equal(runs[0][1], runs[3][1])
