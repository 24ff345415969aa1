"""Pins of the QAM half of the trial chain: the Gray constellations, the max-log demapper and
the complex Gaussian noise of the oracle and of the library's generator tables.

- Points and labels: the OptiCommPy 0.11.0 Gray mappings the reference's own tests compare
  against (tests/oracle/gray_mapping_qam_{4,16,64}.json, used at tests/test_oracle.py:97-105),
  normalised to unit average energy as Constellation::normalized does (constellation.rs:396-398).
- Max-log: the LabelHalves doctest (crates/com-mapping/src/demapper.rs:74-80), and the
  nearest-point form of the max-log LLR, LLR_b = (min_{b=1} |y-p|^2 - min_{b=0} |y-p|^2) / N0,
  which with a uniform prior equals the metric form of demapper.rs:270-297 in exact arithmetic.
- Noise: Rng::complex_gaussian (rng.rs:161-167) and the device generator's Box-Muller, by the
  transform's own identities on the same uniforms and by their moments.
"""
import json
import math
import os

import numpy as np
import pytest

import comfec as F
import oracle as O
from conftest import GOLDEN


def fixture_points(m):
    d = json.load(open(os.path.join(GOLDEN, f"gray_mapping_qam_{m}.json")))
    v = np.array(d["data"]["__array__"]["values"], np.float64)
    p = v[:, 0] + 1j * v[:, 1]
    assert p.size == m
    return p / np.sqrt(np.mean(np.abs(p) ** 2))


def label_bits(m):
    b = int(math.log2(m))
    return np.array([[(i >> (b - 1 - o)) & 1 for o in range(b)] for i in range(m)], np.uint8)


@pytest.mark.parametrize("m", [4, 16, 64])
def test_oracle_points_and_labels_are_opticommpy_gray(m):
    np.testing.assert_allclose(O.qam_points(m), fixture_points(m), rtol=0, atol=1e-12)


@pytest.mark.parametrize("bps", [2, 4, 6])
def test_library_points_and_labels_are_opticommpy_gray(bps):
    np.testing.assert_allclose(F.constellation_points(bps), fixture_points(1 << bps), rtol=0, atol=1e-12)
    with pytest.raises(F.ComfecError) as e:
        F._check(F.lib().cf_constellation_points(bps, np.zeros(3), np.zeros(3), 3))
    assert e.value.kind == "Shape"


def test_gray_neighbours_differ_in_one_bit():
    # the defining property of a Gray mapping: nearest neighbours (distance 2 before the
    # normalisation) carry labels one bit apart
    for m in (4, 16, 64):
        p = fixture_points(m)
        d = np.abs(p[:, None] - p[None, :])
        step = d[d > 1e-9].min()
        for i in range(m):
            for j in np.nonzero(np.abs(d[i] - step) < 1e-9)[0]:
                assert bin(i ^ int(j)).count("1") == 1


def test_label_halves_doctest():
    # demapper.rs:77-79: four labels, bit 1 (the most significant): halves {0, 1} and {2, 3}
    assert O.label_halves_llr([0.0, -1.0, -3.0, -2.0], 1) == 2.0
    # and bit 0: halves {0, 2} (max 0) and {1, 3} (max -1); MSB first
    assert list(O.llr_from_metrics([0.0, -1.0, -3.0, -2.0])) == [2.0, 1.0]


@pytest.mark.parametrize("m,n0", [(4, 0.5), (16, 0.4), (16, 0.05), (64, 0.1)])
def test_max_log_equals_nearest_points(m, n0):
    rng = np.random.default_rng(m)
    p = fixture_points(m)
    y = p[rng.integers(0, m, 300)] + math.sqrt(n0 / 2) * (rng.standard_normal(300) + 1j * rng.standard_normal(300))
    bits = label_bits(m)
    d2 = np.abs(y[:, None] - p[None, :]) ** 2
    want = np.stack([(np.where(bits[:, o] == 1, d2, np.inf).min(1) - np.where(bits[:, o] == 0, d2, np.inf).min(1)) / n0
                     for o in range(bits.shape[1])], 1).ravel()
    got = O.qam_llr(m, y, n0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-11)


def test_complex_gaussian_is_box_muller_on_the_stream():
    # the same stream drawn as uniforms: |g|^2 = -ln(1 - u1) and arg g = 2 pi u2 (rng.rs:161-167)
    g = O.complex_gaussians(7, 2000)
    r = O.Rng(7)
    for k in range(g.size):
        u1, u2 = r.uniform(), r.uniform()
        assert abs(abs(g[k]) ** 2 - (-math.log(1.0 - u1))) <= 1e-12 * max(1.0, abs(g[k]) ** 2)
        ang = math.atan2(g[k].imag, g[k].real) % (2 * math.pi)
        assert min(abs(ang - 2 * math.pi * u2), 2 * math.pi - abs(ang - 2 * math.pi * u2)) < 1e-9


def test_complex_gaussian_moments():
    # variance 1/2 per dimension, zero mean, I and Q uncorrelated (5 sigma bounds)
    n = 200_000
    g = O.complex_gaussians(3, n)
    for part in (g.real, g.imag):
        assert abs(part.mean()) < 5 * math.sqrt(0.5 / n)
        assert abs(part.var() - 0.5) < 5 * 0.5 * math.sqrt(2.0 / n)
    assert abs(np.mean(g.real * g.imag)) < 5 * 0.5 / math.sqrt(n)
    assert abs(np.mean(np.abs(g) ** 2) - 1.0) < 5 * math.sqrt(1.0 / n)


def test_device_noise_is_box_muller_on_philox():
    # or_philox_normals: u1, u2 on (0, 1] from the top 24 bits of Philox words 0 and 1 of
    # counter (index, trial, trial >> 32, point << 4 | 1)
    seed, point, trial = 0x1234_5678_9ABC_DEF0, 3, 2**33 + 5
    for index in range(200):
        g0, g1 = O.philox_normals(seed, point, trial, index)
        w = O.philox([index, trial & 0xFFFFFFFF, trial >> 32, (point << 4) | 1],
                     [seed & 0xFFFFFFFF, seed >> 32])
        u1, u2 = ((w[0] >> 8) + 1) / 2**24, ((w[1] >> 8) + 1) / 2**24
        assert abs(g0 * g0 + g1 * g1 - (-2 * math.log(u1))) <= 1e-12 * max(1.0, -2 * math.log(u1))
        ang = math.atan2(g1, g0) % (2 * math.pi)
        assert min(abs(ang - 2 * math.pi * u2), 2 * math.pi - abs(ang - 2 * math.pi * u2)) < 1e-9
    g = np.array([O.philox_normals(1, 0, t, i) for t in range(20) for i in range(1000)])
    n = g.size
    assert abs(g.mean()) < 5 / math.sqrt(n) and abs(g.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert abs(np.mean(g[:, 0] * g[:, 1])) < 5 / math.sqrt(g.shape[0])


@pytest.mark.parametrize("bps", [1, 2, 4, 6])
def test_philox_trial_noiseless_symbols(bps):
    # at a negligible noise the received symbols are the constellation points of the
    # encoded labels, and the LLR signs are the encoded bits
    code = O.Code.nr(1000, 1800, 0, 0, bps)
    info, y, llr = code.philox_trial(bps, 1e-18, 5, 1, 9)
    assert np.array_equal(info, O.philox_info_bits(5, 1, 9, 1000))
    tx = code.nr_encode(info)
    assert np.array_equal((llr < 0).astype(np.uint8), tx)
    if bps == 1:
        np.testing.assert_allclose(y.real, 1.0 - 2.0 * tx, atol=1e-8)
    else:
        lab = tx.reshape(-1, bps) @ (1 << np.arange(bps - 1, -1, -1))
        np.testing.assert_allclose(y, fixture_points(1 << bps)[lab], atol=1e-8)


def test_oracle_qam_chain_many_equals_single():
    code = O.Code.nr(500, 1200, 0, 0, 4)
    info, llr = code.trial_nr_qam_many(16, 0.3, 11, "qam", 2, 40, 6, threads=3)
    for t in range(6):
        i1, l1 = code.trial_nr_qam(16, 0.3, 11, "qam", 2, 40 + t)
        assert np.array_equal(info[t], i1) and np.array_equal(llr[t], l1)
