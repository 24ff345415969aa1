"""Pins of the oracle's decoder: closed forms on tiny graphs, and the AFF3CT curves.

The AFF3CT runs (validation/aff3ct/*.json, AFF3CT v4.7.1, float64, all-zero codeword)
decode the same PEG matrices with the same rules; the reference accepts a run when the
99 % Wilson intervals of FER overlap (validation/tests/aff3ct.rs:296-333). The decoder
is symmetric, so the all-zero codeword that AFF3CT sends gives the same statistics.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from alist import read_alist
from conftest import GOLDEN


def spc(llr):
    """a single parity check over all the bits"""
    return O.Code.from_rows(len(llr), [list(range(len(llr)))])


def test_min_sum_single_check_closed_form():
    llr = np.array([1.5, -0.25, 2.0, -3.0, 0.75])   # syndrome fails: 2 negatives... make it odd
    llr[0] = -1.5
    code = spc(llr)
    for alpha in (0.75, 1.0):
        post, it = code.decode_llr(llr, rule=O.NMS, alpha=alpha, max_iterations=1)
        a = np.abs(llr)
        for i in range(len(llr)):
            others = np.delete(llr, i)
            want = llr[i] + alpha * np.prod(np.sign(others)) * np.min(np.abs(others))
            assert post[i] == pytest.approx(want, abs=0, rel=1e-15)
    post, _ = code.decode_llr(llr, rule=O.OMS, alpha=0.5, max_iterations=1)
    for i in range(len(llr)):
        others = np.delete(llr, i)
        want = llr[i] + np.prod(np.sign(others)) * max(np.min(np.abs(others)) - 0.5, 0.0)
        assert post[i] == pytest.approx(want, rel=1e-15)
    post, _ = code.decode_llr(llr, rule=O.SPA, max_iterations=1)
    for i in range(len(llr)):
        others = np.delete(llr, i)
        want = llr[i] + 2 * math.atanh(np.prod(np.tanh(others / 2)))
        assert post[i] == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("llr", [[25.0, -30.0, 40.0, 18.0, 22.5], [0.5, -1.25, 3.0, 20.0, 2.0],
                                 [1e-3, 35.0, -36.0, 50.0, 7.0], [-0.0, 3.0, -2.0, 4.0, 5.0]])
def test_f32_sum_product_single_check_closed_form(llr):
    """The FP32 sum-product check (sums of -ln tanh(x/2)) against the reference's f64
    product form, 2 atanh(clamp(prod tanh(l/2))), on one check whose other inputs are past
    the point (|l| > 17) where tanh(l/2) rounds to 1 in FP32 and the product form would
    give atanh(1) = inf; a -0.0 input makes every other message a signed zero."""
    llr = np.array(llr)
    post, it = spc(llr).decode_many_f32(llr[None].astype(np.float32), rule=O.SPA, max_iterations=1)
    assert it[0] == 1
    for i in range(len(llr)):
        p = np.clip(np.prod(np.tanh(np.delete(llr, i) / 2)), -0.999999999999, 0.999999999999)
        assert post[0, i] == pytest.approx(llr[i] + 2 * math.atanh(p), rel=5e-7)
    if llr[0] == 0.0:
        assert np.array_equal(post[0, 1:], llr[1:].astype(np.float32))


@pytest.mark.parametrize("sigma", [0.6, 1.2])
@pytest.mark.parametrize("layered", [True, False], ids=["layered", "flooding"])
def test_f32_sum_product_tracks_f64(sigma, layered):
    """FP32 sum-product on BG2 Z=52 against the f64 reference decoder at noise levels whose
    LLRs and posteriors run far past FP32's tanh saturation: the same iterations and
    decisions on every frame, no NaN, posteriors within 1e-4 relative (measured: <= 8e-6)."""
    ref = O.Code.lifted(2, 52)
    rng = np.random.default_rng(1)
    full = []
    for _ in range(24):
        cw = ref.encode(rng.integers(0, 2, ref.k).astype(np.uint8))
        llr = 2 * (1 - 2.0 * cw + rng.normal(0, sigma, cw.size)) / sigma ** 2
        llr[:104] = 0.0
        full.append(llr)
    full = np.array(full)
    a, ia = ref.decode_many(full, rule=O.SPA, layered=layered, max_iterations=20, threads=8)
    b, ib = ref.decode_many_f32(full.astype(np.float32), rule=O.SPA, layered=layered, max_iterations=20, threads=8)
    assert not np.isnan(b).any() and np.abs(a).max() > 60
    assert np.array_equal(ia, ib) and (ia < 20).all()
    assert np.array_equal(a < 0, b < 0)
    np.testing.assert_allclose(b, a, rtol=1e-4, atol=1e-3)


def test_ties_give_m2_equal_m1_and_minus_zero_is_positive():
    llr = np.array([-0.0, 2.0, -2.0, 2.0, 5.0])  # -0.0 counts positive: one negative -> fails
    code = spc(llr)
    post, it = code.decode_llr(llr, rule=O.NMS, alpha=1.0, max_iterations=1)
    # min over others for bit 0 is 2.0; sign of others is -1
    assert post[0] == -2.0
    # bit 1: others {-0, -2, 2, 5} -> min 0, sign -1 (only -2 negative) -> post = 2 - 0
    assert post[1] == 2.0 and np.signbit(post[1] - 2.0) == False


def layered_numpy(n, rows, llr, alpha, iters):
    post = llr.astype(float).copy()
    c2v = [np.zeros(len(r)) for r in rows]
    def ok(p):
        return all(sum(p[v] < 0 for v in r) % 2 == 0 for r in rows)
    if ok(post):
        return post, 0
    for it in range(1, iters + 1):
        for ri, r in enumerate(rows):
            inc = np.array([post[v] for v in r]) - c2v[ri]
            out = np.zeros(len(r))
            for i in range(len(r)):
                o = np.delete(inc, i)
                s = np.prod(np.where(o < 0, -1.0, 1.0))
                out[i] = s * alpha * np.min(np.abs(o))
            c2v[ri] = out
            for i, v in enumerate(r):
                post[v] = inc[i] + out[i]
        if ok(post):
            return post, it
    return post, iters


def test_layered_hamming_against_brute_restatement():
    # the (7,4) Hamming code of ldpc.rs:1066-1072 through from_rows, against a direct
    # numpy restatement of the layered schedule on random LLRs
    rows = [[0, 1, 2, 4], [0, 1, 3, 5], [0, 2, 3, 6]]
    code = O.Code.from_rows(7, rows)
    rng = np.random.default_rng(4)
    for _ in range(50):
        llr = rng.normal(1.0, 1.5, 7)
        a, ia = code.decode_llr(llr, rule=O.NMS, alpha=0.75, max_iterations=8)
        b, ib = layered_numpy(7, rows, llr, 0.75, 8)
        assert ia == ib
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13)


def test_flooding_two_checks_one_iteration():
    rows = [[0, 1, 2], [1, 2, 3]]
    code = O.Code.from_rows(4, rows)
    llr = np.array([-1.0, 0.5, 2.0, -0.75])
    post, it = code.decode_llr(llr, rule=O.NMS, alpha=1.0, layered=False, max_iterations=1)
    want = llr.copy()
    for r in rows:
        for i, v in enumerate(r):
            o = np.delete(llr[r], i)
            want[v] += np.prod(np.sign(o)) * np.min(np.abs(o))
    np.testing.assert_allclose(post, want, rtol=1e-15)


def test_codeword_input_takes_zero_iterations_and_max_zero_returns_channel():
    code = O.Code.lifted(2, 13)
    cw = code.encode(np.random.default_rng(0).integers(0, 2, code.k).astype(np.uint8))
    llr = 3.0 - 6.0 * cw
    post, it = code.decode_llr(llr, max_iterations=20)
    assert it == 0 and np.array_equal(post, llr)
    noisy = llr + np.random.default_rng(1).normal(0, 3.0, code.n)
    post, it = code.decode_llr(noisy, max_iterations=0)
    assert it == 0 and np.array_equal(post, noisy)


def test_layered_converges_no_slower_than_flooding():
    # ldpc.rs:997-1020 on an NR code
    code = O.Code.lifted(2, 26)
    rng = O.Rng(5)
    u = np.array([rng.bit() for _ in range(code.k)], np.uint8)
    cw = code.encode(u)
    s2 = O.lib.or_bpsk_noise_variance(2.5, code.k / (code.n - 2 * code.z))
    llr = np.zeros(code.n)
    tail = np.zeros(code.n - 2 * code.z)
    O.lib.or_bpsk_awgn_llr(np.ascontiguousarray(cw[2 * code.z:]), len(tail), s2, rng.r, tail)
    llr[2 * code.z:] = tail
    its = {}
    for layered in (False, True):
        for rule, a in ((O.SPA, 0.0), (O.NMS, 0.8), (O.OMS, 0.15)):
            post, it = code.decode_llr(llr, rule=rule, alpha=a, layered=layered, max_iterations=50)
            assert code.is_codeword((post < 0).astype(np.uint8)), (rule, layered)
            assert np.array_equal((post[:code.k] < 0).astype(np.uint8), u)
            its[(rule, layered)] = it
    assert its[(O.SPA, True)] <= its[(O.SPA, False)]


AFF3CT = [
    ("ldpc_peg_n576_k288_layered_nms075_i20", "peg_n576_k288.alist", O.NMS, 0.75, True, 20, [1.5, 2.0, 2.5]),
    ("ldpc_peg_n576_k288_flooding_oms015_i20", "peg_n576_k288.alist", O.OMS, 0.15, False, 20, [1.5, 2.0]),
    ("ldpc_peg_n576_k288_flooding_spa_i50", "peg_n576_k288.alist", O.SPA, 0.0, False, 50, [1.5, 2.0]),
    ("ldpc_peg_n648_k540_layered_spa_i20", "peg_n648_k540.alist", O.SPA, 0.0, True, 20, [3.5, 4.0]),
]


@pytest.mark.parametrize("name,matrix,rule,alpha,layered,iters,points", AFF3CT, ids=[a[0] for a in AFF3CT])
def test_aff3ct_fer_intervals_overlap(name, matrix, rule, alpha, layered, iters, points):
    fx = json.load(open(os.path.join(GOLDEN, "aff3ct_" + name + ".json")))
    n, rows = read_alist(os.path.join(GOLDEN, matrix))
    code = O.Code.from_rows(n, rows)
    k = int(fx["parameters"]["Codec"]["Info. bits (K)"])
    rate = k / n
    seed = O.fnv1a64(name.encode())
    for p in fx["data"]:
        if p["ebn0_db"] not in points:
            continue
        s2 = O.lib.or_bpsk_noise_variance(p["ebn0_db"], rate)
        frames = int(min(12000, max(1500, 60 / p["fer"])))
        llr = np.zeros((frames, n))
        zero = np.zeros(n, np.uint8)
        for t in range(frames):
            r = O.Rng.stream(seed, name, 0, t)
            O.lib.or_bpsk_awgn_llr(zero, n, s2, r.r, llr[t])
        post, _ = code.decode_many(llr, rule=rule, alpha=alpha, layered=layered, max_iterations=iters, threads=8)
        fe = int(np.any(post < 0, axis=1).sum())
        lo, hi = O.wilson(fe, frames, 0.99)
        alo, ahi = O.wilson(p["frame_errors"], p["frames"], 0.99)
        assert not (ahi < lo or hi < alo), (name, p["ebn0_db"], fe / frames, p["fer"])
