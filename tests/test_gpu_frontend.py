"""The reference's Python front end of NrLdpc (bindings/com-py/src/fec/nr.rs:36-140) decoding
on the GPU: the doctest (nr.rs:30-35), the reference's own noiseless decodes of the Sionna
blocks (tests/test_sionna_vectors.py:114-115), and a 1000-block stream decoded in one call
against the oracle's BlockCode::decode_block block by block (f64: bit-identical)."""
import os

import numpy as np
import pytest

import comrs
import oracle as O
from conftest import GOLDEN

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIONNA = dict(np.load(os.path.join(GOLDEN, "sionna_ldpc.npz"), allow_pickle=False))
LDPC = sorted({tuple(int(x) for x in k.split(".")[1][1:].split("_n")) for k in SIONNA if k.endswith(".u")})


def test_doctest_decode():
    print("\n[front end] nr.rs doctest", flush=True)
    code = comrs.NrLdpc(100, 300)
    r = O.Rng(1)
    b = np.array([r.bit() for _ in range(100)], np.uint8)  # comrs.bits(100, comrs.Rng(1))
    assert np.array_equal(code.decode(4.0 - 8.0 * code.encode(b)), b)


@pytest.mark.parametrize("k,n", [kn for kn in LDPC if kn[0] <= 1024])
def test_noiseless_sionna_blocks(k, n):
    u, c = SIONNA[f"ldpc.k{k}_n{n}.u"], SIONNA[f"ldpc.k{k}_n{n}.c"]
    code = comrs.NrLdpc(k, n)
    np.testing.assert_array_equal(code.decode(4.0 - 8.0 * c[0]), u[0])
    # every block of the set in one call
    np.testing.assert_array_equal(code.decode((4.0 - 8.0 * c).ravel()).reshape(u.shape), u)


@pytest.mark.parametrize("rule,alpha", [("normalized_min_sum", 0.75), ("offset_min_sum", 0.15)])
def test_thousand_block_stream_equals_the_oracle(rule, alpha):
    print(f"\n[front end] 1000 blocks, {rule}", flush=True)
    k, e = 1000, 1500
    code = comrs.NrLdpc(k, e, qm=2, rule=rule, alpha=alpha, max_iterations=20)
    ref = O.Code.nr(k, e, 0, 0, 2)
    dec = O.OrDec(O.NMS if rule.startswith("normalized") else O.OMS, alpha, 1, 20)
    s2 = O.lib.or_bpsk_noise_variance(2.0, k / e)
    llrs, want = [], []
    for t in range(1000):
        r = O.Rng.stream(3, "front end", 0, t)
        info = np.array([r.bit() for _ in range(k)], np.uint8)
        llr = np.zeros(e)
        O.lib.or_bpsk_awgn_llr(ref.nr_encode(info), e, s2, r.r, llr)
        llrs.append(llr)
        bits = np.zeros(k, np.uint8)
        O.lib.or_nr_decode_block(ref.h, O.C.byref(dec), llr, bits)
        want.append(bits)
    got = code.decode(np.concatenate(llrs)).reshape(1000, k)
    assert np.array_equal(got, np.array(want))


def test_sum_product_default_against_the_oracle():
    print("\n[front end] default rule (sum-product, 50 it)", flush=True)
    k, e = 500, 1200
    code = comrs.NrLdpc(k, e)
    ref = O.Code.nr(k, e)
    dec = O.OrDec(O.SPA, 0.8, 1, 50)
    s2 = O.lib.or_bpsk_noise_variance(1.5, k / e)
    llrs, want = [], []
    for t in range(64):
        r = O.Rng.stream(4, "spa", 0, t)
        info = np.array([r.bit() for _ in range(k)], np.uint8)
        llr = np.zeros(e)
        O.lib.or_bpsk_awgn_llr(ref.nr_encode(info), e, s2, r.r, llr)
        llrs.append(llr)
        bits = np.zeros(k, np.uint8)
        O.lib.or_nr_decode_block(ref.h, O.C.byref(dec), llr, bits)
        want.append(bits)
    got = code.decode(np.concatenate(llrs)).reshape(64, k)
    assert np.array_equal(got, np.array(want))
