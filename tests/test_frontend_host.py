"""The reference's Python front end of NrLdpc (bindings/com-py/src/fec/nr.rs:36-140) over the
library, on the CPU: construction and its keyword options, the properties, whole-block
encoding against the reference's own test (tests/test_sionna_vectors.py:109-113: every
Sionna vector block through one encode call), and the errors the binding raises."""
import os

import numpy as np
import pytest

import comrs
import oracle as O
from conftest import GOLDEN

SIONNA = dict(np.load(os.path.join(GOLDEN, "sionna_ldpc.npz"), allow_pickle=False))
LDPC = sorted({tuple(int(x) for x in k.split(".")[1][1:].split("_n")) for k in SIONNA if k.endswith(".u")})


def test_doctest_construction():
    # nr.rs:30-33
    code = comrs.NrLdpc(100, 300)
    assert (code.base_graph, code.lifting_size) == (2, 18)
    assert (code.k, code.n) == (100, 300) and code.rate == pytest.approx(100 / 300)
    assert repr(code) == "NrLdpc(k=100, n=300, base_graph=2, lifting_size=18)"


@pytest.mark.parametrize("k,n", LDPC)
def test_whole_block_encode_matches_sionna(k, n):
    u, c = SIONNA[f"ldpc.k{k}_n{n}.u"], SIONNA[f"ldpc.k{k}_n{n}.c"]
    code = comrs.NrLdpc(k, n)
    np.testing.assert_array_equal(code.encode(u.ravel()).reshape(c.shape), c)


def test_keyword_options_reach_the_code():
    code = comrs.NrLdpc(500, 1200, base_graph=1, rv=2, qm=4, rule="offset_min_sum", alpha=0.15,
                        max_iterations=12, layered=False)
    ref = O.Code.nr(500, 1200, 1, 2, 4)
    assert (code.base_graph, code.lifting_size) == (ref.bg, ref.z)
    info = np.random.default_rng(3).integers(0, 2, 3 * 500).astype(np.uint8)
    want = np.concatenate([ref.nr_encode(info[i * 500:(i + 1) * 500]) for i in range(3)])
    np.testing.assert_array_equal(code.encode(info), want)


def test_errors_are_the_bindings():
    code = comrs.NrLdpc(100, 300)
    with pytest.raises(ValueError, match=r"^inconsistent shapes: 301 LLRs is not a multiple of the block length 300$"):
        code.decode(np.zeros(301))
    with pytest.raises(ValueError, match=r"^inconsistent shapes: 150 bits is not a multiple of the block length 100$"):
        code.encode(np.zeros(150, np.uint8))
    with pytest.raises(ValueError, match="bits must be 0 or 1"):
        code.encode(np.full(100, 2))
    with pytest.raises(ValueError, match="^invalid configuration: rule must be sum_product, min_sum"):
        comrs.NrLdpc(100, 300, rule="belief")
    with pytest.raises(ValueError, match="^invalid configuration: "):
        comrs.NrLdpc(9000, 10000)
    with pytest.raises(TypeError):
        comrs.NrLdpc(100, 300, 2)  # the options are keyword-only
    assert code.encode(np.zeros(0, np.uint8)).size == 0
