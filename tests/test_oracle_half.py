"""Pins of the oracle's binary16-storage decoder (the target of the device FP16 mode).

- The conversions: binary32 -> binary16 round-to-nearest-even against numpy's float16
  cast on random values of every magnitude, exact ties and the special cases;
  binary16 -> binary32 on all 65 536 encodings.
- The decoder itself: where no rounding can happen (integer LLRs, offset min-sum with
  beta = 1, every value an integer below 2048, exactly representable in binary16) it must
  be the reference's f64 algorithm bit for bit — same schedule, rule, stop and iteration
  count — so the f64 oracle (pinned by the reference's vectors) pins it.
- Where rounding happens it does happen (posteriors differ from FP32), and the decisions
  stay those of the FP32 decoder on nearly every frame.
"""
import numpy as np
import pytest

import oracle as O


def test_f32_to_f16_matches_numpy():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 6, 20000)).astype(np.float32)
    h = rng.integers(0, 0x7bff, 4000).astype(np.uint16)
    lo = h.view(np.float16).astype(np.float64)
    hi = (h + 1).astype(np.uint16).view(np.float16).astype(np.float64)
    ties = ((lo + hi) / 2).astype(np.float32)  # exactly halfway: ties go to even
    assert np.all(ties.astype(np.float64) == (lo + hi) / 2)
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 1e-8, 2.0 ** -25, 2.0 ** -24,
                        1.5 * 2.0 ** -24, 2.5 * 2.0 ** -24, 2.0 ** -14, np.inf, -np.inf], np.float32)
    x = np.concatenate([x, ties, -ties, special])
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(O.f32_to_f16_bits(x), want)


def test_f16_to_f32_all_encodings():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = O.f16_bits_to_f32(h)
    want = h.view(np.float16).astype(np.float32)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all()


@pytest.mark.parametrize("bg,z", [(2, 52), (1, 16), (1, 40)])
def test_exact_integer_case_equals_the_f64_decoder(bg, z):
    ref = O.Code.lifted(bg, z)
    rng = np.random.default_rng(bg * 100 + z)
    count = 24
    cws = []
    for _ in range(count):
        u = rng.integers(0, 2, ref.k).astype(np.uint8)
        cws.append(ref.encode(u))
    cws = np.array(cws)
    # integer LLRs: mostly right, some wrong, some erased
    mag = rng.integers(0, 6, cws.shape)
    flip = rng.random(cws.shape) < 0.03
    llr = np.where(cws == 1, -1, 1) * mag * np.where(flip, -1, 1)
    llr[:, :2 * z] = 0
    full = llr.astype(np.float64)
    seen = set()
    for lim in (1, 3, 6):
        post64, it64 = ref.decode_many(full, rule=O.OMS, alpha=1.0, layered=True, max_iterations=lim, threads=8)
        assert np.all(np.abs(post64) < 2048) and np.all(post64 == np.round(post64))
        post16, it16 = ref.decode_many_f16(full.astype(np.float32), rule=O.OMS, alpha=1.0, max_iterations=lim,
                                           threads=8)
        assert np.array_equal(it16, it64)
        assert np.array_equal(post16.astype(np.float64), post64 + 0.0)  # +0.0: a zero is stored as +0.0
        seen |= set((it64 < lim).tolist())
    assert seen == {True, False}  # some frames converge, some run to the limit


def test_rounding_happens_and_decisions_track_fp32():
    ref = O.Code.lifted(2, 52)
    fulls = []
    for t in range(64):
        _, llr = ref.trial_bpsk_lifted(1.5, 1, "half", 0, t)
        fulls.append(llr)
    full = np.array(fulls).astype(np.float32)
    p32, i32 = ref.decode_many_f32(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=8)
    p16, i16 = ref.decode_many_f16(full, rule=O.NMS, alpha=0.75, max_iterations=20, threads=8)
    assert not np.array_equal(p16, p32)
    # every stored posterior is a binary16 value
    assert np.array_equal(p16.astype(np.float16).astype(np.float32), p16)
    same = ((p16 < 0) == (p32 < 0)).all(1)
    assert same.mean() >= 0.9
