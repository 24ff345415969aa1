"""Pins of the CPU oracle against reference-held vectors and closed forms.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a
wrong index or shift, a transposed table column) fails it.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

SIONNA = dict(np.load(os.path.join(GOLDEN, "sionna_ldpc.npz"), allow_pickle=False))
PAIRS = sorted({tuple(int(x) for x in k.split(".")[1][1:].split("_n")) for k in SIONNA})


def test_rng_reference_values():
    # crates/com-core/src/rng.rs:300-305 (xoshiro256** seeded by SplitMix64 from 0)
    assert O.Rng(0).next_u64() == 0x99EC5F36CB75F2B4
    assert O.fnv1a64(b"a") == 0xAF63DC4C8601EC8C


def test_rng_published_splitmix_and_moments():
    # SplitMix64 from state 0: first output of Vigna's reference C code
    import ctypes
    s = ctypes.c_uint64(0)
    assert O.lib.or_splitmix64(ctypes.byref(s)) == 0xE220A8397B1DCDAF
    # uniform has 53 bits; bit is the top bit; gaussian moments (rng.rs:317-331)
    r = O.Rng(7)
    g = np.array([r.gaussian() for _ in range(100_000)])
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.02
    r1, r2 = O.Rng(11), O.Rng(11)
    for _ in range(100):
        assert r1.uniform() == (r2.next_u64() >> 11) * 2.0 ** -53
    r1, r2 = O.Rng(12), O.Rng(12)
    assert all(r1.bit() == r2.next_u64() >> 63 for _ in range(100))


def test_stream_keying():
    # rng.rs:270-278 restated in Python from the primitives
    def sm(st):
        st = (st + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)
    base, name, cond, trial = 7, "nr", 3, 12345
    c = sm(sm(sm(base ^ O.fnv1a64(name.encode())) ^ cond) ^ trial)
    assert O.Rng.stream(base, name, cond, trial).next_u64() == O.Rng(c).next_u64()
    assert O.Rng.stream(base, name, cond, trial).next_u64() != O.Rng.stream(base, name, cond, trial + 1).next_u64()


@pytest.mark.parametrize("k,n", PAIRS)
def test_sionna_encoder_vectors(k, n):
    # tests/test_sionna_vectors.py:104-115: NrLdpc(k, n).encode is bit-exact, and the
    # noiseless LLRs 4 - 8c decode to u for K' <= 1024
    u, c = SIONNA[f"ldpc.k{k}_n{n}.u"], SIONNA[f"ldpc.k{k}_n{n}.c"]
    code = O.Code.nr(k, n)
    for i in range(len(u)):
        assert np.array_equal(code.nr_encode(u[i]), c[i])
    if k <= 1024:
        bits, it = code.decode_block(4.0 - 8.0 * c[0].astype(np.float64), rule=O.SPA, max_iterations=50)
        assert np.array_equal(bits, u[0])


def test_doctest_nrldpc_100_300():
    # crates/com-fec/src/nr/ldpc.rs:33-43
    code = O.Code.nr(100, 300)
    assert (code.bg, code.z) == (2, 18)
    info = np.array([(i % 3 == 0) for i in range(100)], np.uint8)
    c = code.nr_encode(info)
    llr = np.where(c == 0, 4.0, -4.0)
    bits, _ = code.decode_block(llr, rule=O.SPA, max_iterations=50)
    assert np.array_equal(bits, info)


@pytest.mark.parametrize("bg", [1, 2])
def test_every_small_lifting_size_encodes_codewords(bg):
    # nr/ldpc.rs:265-282, plus the table structure of :249-262 through H c = 0
    rng = np.random.default_rng(bg)
    for z in [2, 3, 5, 7, 9, 11, 13, 15, 16, 26, 52, 64]:
        code = O.Code.lifted(bg, z)
        assert code.n == (68 if bg == 1 else 52) * z
        assert code.edges == (316 if bg == 1 else 197) * z
        u = rng.integers(0, 2, code.k).astype(np.uint8)
        assert code.is_codeword(code.encode(u))


def test_big_codes_encode_codewords():
    rng = np.random.default_rng(3)
    for bg, z in [(1, 384), (2, 384), (1, 352), (2, 208)]:
        code = O.Code.lifted(bg, z)
        u = rng.integers(0, 2, code.k).astype(np.uint8)
        assert code.is_codeword(code.encode(u))


def test_rate_matching_formula():
    # nr/ldpc.rs:287-330: the selection equals the index formula, fillers carry 30.0
    for k_info, e, rv, qm in [(100, 300, 0, 1), (500, 1200, 2, 4), (1000, 1500, 3, 6), (60, 1000, 1, 2)]:
        code = O.Code.nr(k_info, e, 0, rv, qm)
        z, k = code.z, code.k
        ncb = code.n - 2 * z
        num, den = {(0,): (0, 1)}.get((rv,), None) or {
            (1, 1): (17, 66), (1, 2): (33, 66), (1, 3): (56, 66),
            (2, 1): (13, 50), (2, 2): (25, 50), (2, 3): (43, 50)}[(code.bg, rv)]
        k0 = num * ncb // (den * z) * z
        sel, j = [], 0
        while len(sel) < e:
            col = 2 * z + (k0 + j) % ncb
            if not (k_info <= col < k):
                sel.append(col)
            j += 1
        want = [sel[(t % qm) * (e // qm) + t // qm] for t in range(e)]
        assert list(code.bit_map()) == want
        llr = np.random.default_rng(21).standard_normal(e)
        full = np.zeros(code.n)
        for i, l in zip(want, llr):
            full[i] += l
        full[k_info:k] = 30.0
        assert np.array_equal(code.full_llr(llr), full)


def test_wilson_and_z():
    # stats.rs test: 0 errors in 1000 at 0.95 -> (0, ~3.83e-3); z_score(0.95) ~ 1.959964
    lo, hi = O.wilson(0, 1000, 0.95)
    assert lo == 0.0 and abs(hi - 3.83e-3) < 1e-4
    assert abs(O.lib.or_z_score(0.95) - 1.959964) < 1e-5
    assert abs(O.lib.or_qfuncinv(0.025) - 1.959964) < 1e-6
    # erfc against scipy (a different implementation of the same function)
    from scipy.special import erfc
    for x in [-3.0, -0.3, 0.0, 0.2, 0.46875, 0.5, 1.7, 3.9, 4.1, 8.0, 27.0]:
        assert abs(O.lib.or_erfc(x) - erfc(x)) <= 1e-15 + 1e-14 * erfc(x)
    # Wilson against its closed form at a non-trivial point
    k, n, z = 37, 1000, O.lib.or_z_score(0.95)
    p = k / n
    c = (p + z * z / (2 * n)) / (1 + z * z / n)
    h = z * np.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / (1 + z * z / n)
    assert np.allclose(O.wilson(k, n), (c - h, c + h), rtol=1e-15)


def test_philox_known_answers():
    # Random123 known-answer vectors of philox4x32-10 (Salmon et al., SC'11)
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
