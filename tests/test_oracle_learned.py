"""Pins of the oracle's learned offset/scale min-sum (no reference counterpart).

- with a = alpha, b = 0 it is the reference's normalized min-sum bit for bit, and with
  a = 1, b = beta its offset min-sum (the decode of ldpc.rs:509-531, restated separately);
- its hand-written backward equals the central finite differences of its loss.
"""
import numpy as np
import pytest

import oracle as O


def noisy(code, ebn0, seed):
    rng = np.random.default_rng(seed)
    cw = code.encode(rng.integers(0, 2, code.k).astype(np.uint8))
    s2 = O.lib.or_bpsk_noise_variance(ebn0, code.k / (code.n - 2 * code.z))
    llr = 2 * ((1 - 2.0 * cw) + np.sqrt(s2) * rng.standard_normal(code.n)) / s2
    llr[:2 * code.z] = 0.0
    return cw, llr


@pytest.mark.parametrize("bg,z", [(2, 13), (1, 7), (2, 52)])
def test_reduces_to_the_plain_rules_bit_for_bit(bg, z):
    code = O.Code.lifted(bg, z)
    rows = 46 if bg == 1 else 42
    for seed in range(6):
        cw, llr = noisy(code, 1.0, seed)
        post_ref, it_ref = code.decode_llr(llr, rule=O.NMS, alpha=0.75, max_iterations=12)
        post, it = code.learned_decode(llr, np.full((12, rows), 0.75), np.zeros((12, rows)))
        assert it == it_ref and np.array_equal(post, post_ref)
        post_ref, it_ref = code.decode_llr(llr, rule=O.OMS, alpha=0.15, max_iterations=12)
        post, it = code.learned_decode(llr, np.ones((12, rows)), np.full((12, rows), 0.15))
        assert it == it_ref and np.array_equal(post, post_ref)


@pytest.mark.parametrize("bg,z,T", [(2, 5, 4), (1, 3, 3), (2, 13, 5)])
def test_backward_matches_central_finite_differences(bg, z, T):
    code = O.Code.lifted(bg, z)
    rows = 46 if bg == 1 else 42
    rng = np.random.default_rng(bg * 100 + z)
    cw, llr = noisy(code, 0.5, z)
    a = 0.75 + 0.1 * rng.standard_normal((T, rows))
    b = 0.05 + 0.1 * np.abs(rng.standard_normal((T, rows)))
    loss, ga, gb = code.learned_loss_grad(llr, cw, a, b)
    assert loss > 0
    h = 1e-6
    checked = 0
    for t in range(T):
        for i in rng.choice(rows, 8, replace=False):
            for P, G in ((a, ga), (b, gb)):
                P[t, i] += h
                lp, _, _ = code.learned_loss_grad(llr, cw, a, b)
                P[t, i] -= 2 * h
                lm, _, _ = code.learned_loss_grad(llr, cw, a, b)
                P[t, i] += h
                fd = (lp - lm) / (2 * h)
                assert abs(fd - G[t, i]) <= 1e-6 * abs(G[t, i]) + 1e-11, (t, i, fd, G[t, i])
                checked += abs(G[t, i]) > 1e-9
    assert checked >= 10
