"""Codes given by their parity-check rows on the GPU (the general kernel), through the C ABI:
the AFF3CT fixture configurations of validation/aff3ct/*.json (PEG matrices, BPSK over AWGN,
the all-zero codeword AFF3CT sends) replayed on the device.

- f64 messages: posteriors and iteration counts bit-identical to the oracle (layered NMS,
  flooding OMS; sum-product within the device's tanh/atanh rounding);
- FP32 messages: bit-identical to the FP32 oracle;
- frame error rates of the device decoder inside the 99 % Wilson interval of AFF3CT's
  (the acceptance rule of validation/tests/aff3ct.rs:296-333);
- the host entry points return the information positions (decode_block, app_into).
"""
import json
import os

import numpy as np
import pytest

import comfec as F
import oracle as O
from alist import read_alist
from conftest import GOLDEN
from gf2 import nonpivot_columns

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

AFF3CT = [
    ("ldpc_peg_n576_k288_layered_nms075_i20", "peg_n576_k288.alist", O.NMS, "normalized_min_sum", 0.75, True, 20),
    ("ldpc_peg_n576_k288_flooding_oms015_i20", "peg_n576_k288.alist", O.OMS, "offset_min_sum", 0.15, False, 20),
    ("ldpc_peg_n576_k288_flooding_spa_i50", "peg_n576_k288.alist", O.SPA, "sum_product", 0.0, False, 50),
    ("ldpc_peg_n648_k540_layered_spa_i20", "peg_n648_k540.alist", O.SPA, "sum_product", 0.0, True, 20),
]


def zero_codeword_llrs(name, n, ebn0_db, rate, frames, cond=0):
    """The oracle's BPSK/AWGN chain on the all-zero codeword, stream (name, cond, t)."""
    s2 = O.lib.or_bpsk_noise_variance(ebn0_db, rate)
    seed = O.fnv1a64(name.encode())
    llr = np.zeros((frames, n))
    zero = np.zeros(n, np.uint8)
    for t in range(frames):
        r = O.Rng.stream(seed, name, cond, t)
        O.lib.or_bpsk_awgn_llr(zero, n, s2, r.r, llr[t])
    return llr


def setup(matrix, name, alpha, layered, iters):
    n, rows = read_alist(os.path.join(GOLDEN, matrix))
    code = F.NrLdpc.from_rows(n, rows).set_decoder(name, alpha, layered, iters)
    return n, rows, code, O.Code.from_rows(n, rows)


@pytest.mark.parametrize("cfg", AFF3CT, ids=[a[0] for a in AFF3CT])
def test_device_matches_the_oracle(cfg):
    fx_name, matrix, orule, name, alpha, layered, iters = cfg
    print(f"\n[rows] {fx_name}: device vs oracle", flush=True)
    n, rows, code, ref = setup(matrix, name, alpha, layered, iters)
    fx = json.load(open(os.path.join(GOLDEN, "aff3ct_" + fx_name + ".json")))
    k = int(fx["parameters"]["Codec"]["Info. bits (K)"])
    ebn0 = fx["data"][len(fx["data"]) // 2]["ebn0_db"]
    llr = zero_codeword_llrs(fx_name, n, ebn0, k / n, 128, cond=7)
    post_ref, it_ref = ref.decode_many(llr, rule=orule, alpha=alpha, layered=layered, max_iterations=iters, threads=16)
    bits, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    post = post.cpu().numpy()
    assert np.array_equal(its.cpu().numpy(), it_ref)
    if orule == O.SPA:
        # device tanh/atanh against glibc's: decisions equal, posteriors within 1e-9 relative
        assert np.array_equal(post < 0, post_ref < 0)
        assert np.max(np.abs(post - post_ref) / np.maximum(np.abs(post_ref), 1.0)) < 1e-9
    else:
        assert np.array_equal(post, post_ref)
        p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=orule, alpha=alpha, layered=layered,
                                       max_iterations=iters, threads=16)
        _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
        assert np.array_equal(its32.cpu().numpy(), i32)
        assert np.array_equal(post32.cpu().numpy(), p32)
    assert len(set(it_ref.tolist())) >= 3


@pytest.mark.parametrize("cfg", AFF3CT, ids=[a[0] for a in AFF3CT])
def test_device_fer_inside_aff3ct_intervals(cfg):
    fx_name, matrix, orule, name, alpha, layered, iters = cfg
    print(f"\n[rows] {fx_name}: FER vs AFF3CT", flush=True)
    n, rows, code, ref = setup(matrix, name, alpha, layered, iters)
    fx = json.load(open(os.path.join(GOLDEN, "aff3ct_" + fx_name + ".json")))
    k = int(fx["parameters"]["Codec"]["Info. bits (K)"])
    checked = 0
    for ci, p in enumerate(fx["data"]):
        if p["fer"] < 2e-3 or p["fer"] > 0.9:
            continue
        frames = int(min(20000, max(2000, 200 / p["fer"])))
        llr = zero_codeword_llrs(fx_name, n, p["ebn0_db"], k / n, frames, cond=ci)
        _, _, post = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
        fe = int((post < 0).any(1).sum())
        lo, hi = O.wilson(fe, frames, 0.99)
        alo, ahi = O.wilson(p["frame_errors"], p["frames"], 0.99)
        print(f"  {p['ebn0_db']} dB: device {fe}/{frames} = {fe / frames:.4g}, AFF3CT {p['fer']:.4g}", flush=True)
        assert not (ahi < lo or hi < alo), (fx_name, p["ebn0_db"], fe / frames, p["fer"])
        checked += 1
    assert checked >= 2


def test_host_entry_points_return_the_information_positions():
    print("\n[rows] decode_block / app_into", flush=True)
    n, rows, code, ref = setup("peg_n576_k288.alist", "normalized_min_sum", 0.75, True, 20)
    info = np.array(nonpivot_columns(n, rows))
    rng = np.random.default_rng(4)
    u = rng.integers(0, 2, code.k).astype(np.uint8)
    cw = code.encode_block(u)
    s2 = O.lib.or_bpsk_noise_variance(2.5, code.k / n)
    llr = 2.0 * ((1.0 - 2.0 * cw) + rng.normal(0.0, np.sqrt(s2), n)) / s2
    post_ref, it_ref = ref.decode_llr(llr, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20)
    assert np.array_equal(code.decode_block(llr), (post_ref[info] < 0).astype(np.uint8))
    apriori = rng.normal(0.0, 2.0, code.k)
    shifted = llr.copy()
    shifted[info] += apriori
    post2, it2 = ref.decode_llr(shifted, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20)
    out, it = code.app(llr, apriori)
    assert it == it2 and np.array_equal(out, post2[info])


DVBS2 = "ldpc_dvbs2_short_r1_2_layered_spa_i50"


@pytest.fixture(scope="module")
def dvbs2():
    """The DVB-S2 short rate-1/2 code (built once: the host's reduced form takes seconds)."""
    from dvbs2 import dvbs2_short_r1_2_rows
    n, k, rows = dvbs2_short_r1_2_rows()
    code = F.NrLdpc.from_rows(n, rows).set_decoder("sum_product", 0.0, True, 50)
    return n, k, rows, code


def test_dvbs2_device_matches_the_oracle(dvbs2):
    """Layered sum-product on the DVB-S2 code (one lane per codeword walking the 9000
    checks in order): f64 against the oracle (iterations and decisions identical,
    posteriors within the device's tanh/atanh rounding), FP32 against the FP32 oracle
    (iterations and decisions identical, posteriors within libm rounding)."""
    print(f"\n[rows] {DVBS2}: device vs oracle", flush=True)
    n, k, rows, code = dvbs2
    ref = O.Code.from_rows(n, rows)
    llr = zero_codeword_llrs(DVBS2, n, 0.8, k / n, 24, cond=7)
    post_ref, it_ref = ref.decode_many(llr, rule=O.SPA, layered=True, max_iterations=50, threads=16)
    bits, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    post = post.cpu().numpy()
    assert np.array_equal(its.cpu().numpy(), it_ref) and len(set(it_ref.tolist())) >= 3
    assert np.array_equal(post < 0, post_ref < 0)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :k] < 0).astype(np.uint8))
    assert np.max(np.abs(post - post_ref) / np.maximum(np.abs(post_ref), 1.0)) < 1e-9
    p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=O.SPA, layered=True, max_iterations=50, threads=16)
    _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
    post32 = post32.cpu().numpy()
    assert np.array_equal(its32.cpu().numpy(), i32) and np.array_equal(post32 < 0, p32 < 0)
    # device exp/expm1/log1p against glibc's over up to 50 sweeps of 9000 checks: measured
    # 1 posterior in 388 800 off by 1.1e-3 (at 0.13), the rest within 1e-4 relative
    np.testing.assert_allclose(post32, p32, rtol=1e-4, atol=1e-2)


def test_dvbs2_fer_inside_aff3ct_intervals(dvbs2):
    """The AFF3CT run of the DVB-S2 code (validation/aff3ct/ldpc_dvbs2_short_r1_2_layered_spa_i50.json:
    BPSK over AWGN, the all-zero codeword, layered sum-product, 50 iterations) replayed on
    the device with FP32 messages: frame errors over the coded bits inside AFF3CT's
    99 % Wilson interval at every point with a frame error rate in [2e-3, 0.9] (the
    acceptance rule of validation/tests/aff3ct.rs:296-333). Channel noise from torch's
    generator on the device, seeded per point."""
    n, k, rows, code = dvbs2
    fx = json.load(open(os.path.join(GOLDEN, "aff3ct_" + DVBS2 + ".json")))
    checked = 0
    for ci, p in enumerate(fx["data"]):
        if p["fer"] < 2e-3 or p["fer"] > 0.9:
            continue
        frames = int(min(20000, max(2000, 200 / p["fer"])))
        s2 = O.lib.or_bpsk_noise_variance(p["ebn0_db"], k / n)
        g = torch.Generator(device="cuda")
        g.manual_seed(1000 + ci)
        noise = torch.randn((frames, n), generator=g, dtype=torch.float32, device="cuda")
        llr = ((1.0 + float(np.sqrt(s2)) * noise) * float(2.0 / s2)).contiguous()
        print(f"  {p['ebn0_db']} dB: {frames} frames", flush=True)
        # frame errors over all N coded bits, as AFF3CT and the reference count them
        # (validation/tests/aff3ct.rs:151-158, bits_per_frame = N)
        _, its, post = code.decode_batch(llr, want_post=True, want_bits=False)
        fe = int((post < 0).any(1).sum().item())
        del post
        lo, hi = O.wilson(fe, frames, 0.99)
        alo, ahi = O.wilson(p["frame_errors"], p["frames"], 0.99)
        print(f"  {p['ebn0_db']} dB: device {fe}/{frames} = {fe / frames:.4g}, AFF3CT {p['fer']:.4g}, "
              f"mean iterations {its.double().mean().item():.2f}", flush=True)
        assert not (ahi < lo or hi < alo), (DVBS2, p["ebn0_db"], fe / frames, p["fer"])
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("rule", [(O.OMS, "offset_min_sum", 0.15), (O.NMS, "normalized_min_sum", 1.0)],
                         ids=["oms015", "min_sum"])
def test_layered_groups_other_rules_bit_exact(rule):
    """The group kernel's other min-sum rules on the PEG (576, 288) code (checks of 5 to 7
    edges, a group of 8 lanes a codeword), layered, 20 iterations: f64 and FP32 posteriors
    and iteration counts bit-identical to the oracle's, on more codewords than groups fit
    in one wavefront and with a batch that is not a multiple of 8."""
    orule, name, alpha = rule
    print(f"\n[rows] layered {name} {alpha}", flush=True)
    n, rows, code, ref = setup("peg_n576_k288.alist", name, alpha, True, 20)
    llr = zero_codeword_llrs("groups_" + name, n, 2.5, 0.5, 203, cond=3)
    post_ref, it_ref = ref.decode_many(llr, rule=orule, alpha=alpha, layered=True, max_iterations=20, threads=16)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref) and np.array_equal(post.cpu().numpy(), post_ref)
    p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=orule, alpha=alpha, layered=True,
                                   max_iterations=20, threads=16)
    _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
    assert np.array_equal(its32.cpu().numpy(), i32) and np.array_equal(post32.cpu().numpy(), p32)
    assert len(set(it_ref.tolist())) >= 3


def test_layered_groups_sum_product_f32():
    """FP32 layered sum-product on the PEG (576, 288) code through the group kernel (prefix
    and suffix sums by lane shifts): the FP32 oracle's iterations and decisions on every
    frame; the posteriors differ only by the last bits of device exp/expm1/log1p against
    glibc's, carried through up to 30 iterations: the median difference is below 1e-6
    relative, the largest measured 0.019 (on posteriors of magnitude 10 to 20), 1e-3
    relative or 0.05 is the bound."""
    print("\n[rows] layered sum-product FP32", flush=True)
    n, rows, code, ref = setup("peg_n576_k288.alist", "sum_product", 0.0, True, 30)
    llr = zero_codeword_llrs("groups_spa", n, 2.0, 0.5, 96, cond=4).astype(np.float32)
    p32, i32 = ref.decode_many_f32(llr, rule=O.SPA, layered=True, max_iterations=30, threads=16)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    post = post.cpu().numpy()
    assert np.array_equal(its.cpu().numpy(), i32) and np.array_equal(post < 0, p32 < 0)
    rel = np.abs(post - p32) / np.maximum(np.abs(p32), 1.0)
    assert np.median(rel) < 1e-6
    np.testing.assert_allclose(post, p32, rtol=1e-3, atol=5e-2)


def test_layered_groups_of_32_bit_exact():
    """Checks of 17 to 19 edges (the PEG (648, 540) code) take a group of 32 lanes, whose
    exchanges cross a DPP row through ds_swizzle: layered normalized min-sum f64 and FP32
    posteriors and iteration counts bit-identical to the oracle's."""
    print("\n[rows] groups of 32, layered NMS", flush=True)
    n, rows, code, ref = setup("peg_n648_k540.alist", "normalized_min_sum", 0.75, True, 20)
    llr = zero_codeword_llrs("groups32", n, 3.5, 540 / 648, 101, cond=5)
    post_ref, it_ref = ref.decode_many(llr, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=16)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref) and np.array_equal(post.cpu().numpy(), post_ref)
    p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=O.NMS, alpha=0.75, layered=True,
                                   max_iterations=20, threads=16)
    _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
    assert np.array_equal(its32.cpu().numpy(), i32) and np.array_equal(post32.cpu().numpy(), p32)
    assert len(set(it_ref.tolist())) >= 3


def test_groups_edge_cases():
    """The group kernel with an empty batch, a limit of 0 iterations (the channel LLRs come
    back, 0 iterations), a codeword given as input (0 iterations) and a batch of one."""
    print("\n[rows] group kernel edge cases", flush=True)
    n, rows, code, ref = setup("peg_n576_k288.alist", "normalized_min_sum", 0.75, True, 0)
    bits, its, post = code.decode_batch(torch.zeros((0, n), dtype=torch.float64, device="cuda"), want_post=True)
    assert bits.shape == (0, code.k) and its.shape == (0,) and post.shape == (0, n)
    llr = zero_codeword_llrs("groups_edge", n, 0.5, 0.5, 5, cond=6)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    assert (its.cpu().numpy() == 0).all() and np.array_equal(post.cpu().numpy(), llr)
    code.set_decoder("normalized_min_sum", 0.75, True, 20)
    clean = np.full((3, n), 4.0)
    _, its, post = code.decode_batch(torch.from_numpy(clean).cuda(), want_post=True)
    assert (its.cpu().numpy() == 0).all() and np.array_equal(post.cpu().numpy(), clean)
    post_ref, it_ref = ref.decode_llr(llr[1], rule=O.NMS, alpha=0.75, layered=True, max_iterations=20)
    _, its, post = code.decode_batch(torch.from_numpy(llr[1:2].copy()).cuda(), want_post=True)
    assert its.item() == it_ref and np.array_equal(post.cpu().numpy()[0], post_ref)


def test_checks_listing_a_variable_twice():
    """A check may list a variable twice (the reference's from_rows takes such rows): the
    device then stores a check's posteriors edge after edge, so that the later edge's value
    stays, as in the oracle. The PEG (576, 288) code with the first variable of every
    seventh check listed again at its end, layered NMS 0.75, 20 iterations: f64 and FP32
    posteriors and iteration counts bit-identical to the oracle's."""
    print("\n[rows] checks listing a variable twice", flush=True)
    n, rows = read_alist(os.path.join(GOLDEN, "peg_n576_k288.alist"))
    rows = [list(r) + ([r[0]] if i % 7 == 0 else []) for i, r in enumerate(rows)]
    code = F.NrLdpc.from_rows(n, rows).set_decoder("normalized_min_sum", 0.75, True, 20)
    ref = O.Code.from_rows(n, rows)
    llr = zero_codeword_llrs("twice", n, 2.5, 0.5, 70, cond=7)
    post_ref, it_ref = ref.decode_many(llr, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=16)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref) and np.array_equal(post.cpu().numpy(), post_ref)
    p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=O.NMS, alpha=0.75, layered=True,
                                   max_iterations=20, threads=16)
    _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
    assert np.array_equal(its32.cpu().numpy(), i32) and np.array_equal(post32.cpu().numpy(), p32)
    assert len(set(it_ref.tolist())) >= 3


def test_layered_groups_of_16_bit_exact():
    """Checks of 9 to 16 edges take a group of 16 lanes (the row mirror between the two
    halves): the PEG (576, 288) checks joined in pairs (check 2j with check 2j + 1, a
    variable both list kept once; 144 checks of 10 to 13 edges; Eb/N0 2.5 dB at rate 3/4), layered normalized min-sum
    f64 and FP32 posteriors and iteration counts bit-identical to the oracle's."""
    print("\n[rows] groups of 16, layered NMS", flush=True)
    n, rows = read_alist(os.path.join(GOLDEN, "peg_n576_k288.alist"))
    joined = [list(dict.fromkeys(list(rows[2 * j]) + list(rows[2 * j + 1]))) for j in range(len(rows) // 2)]
    degs = [len(r) for r in joined]
    assert 8 < min(degs) and max(degs) <= 16
    code = F.NrLdpc.from_rows(n, joined).set_decoder("normalized_min_sum", 0.75, True, 20)
    ref = O.Code.from_rows(n, joined)
    llr = zero_codeword_llrs("groups16", n, 2.5, 0.75, 90, cond=8)
    post_ref, it_ref = ref.decode_many(llr, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=16)
    _, its, post = code.decode_batch(torch.from_numpy(llr).cuda(), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref) and np.array_equal(post.cpu().numpy(), post_ref)
    p32, i32 = ref.decode_many_f32(llr.astype(np.float32), rule=O.NMS, alpha=0.75, layered=True,
                                   max_iterations=20, threads=16)
    _, its32, post32 = code.decode_batch(torch.from_numpy(llr.astype(np.float32)).cuda(), want_post=True)
    assert np.array_equal(its32.cpu().numpy(), i32) and np.array_equal(post32.cpu().numpy(), p32)
    assert len(set(it_ref.tolist())) >= 3
