"""CPU checks of the C ABI: the library loads, exports every entry point its header
declares, and its host logic (construction, rate matching, encoder, errors) matches the
oracle and the reference-held Sionna vectors. Nothing here launches a kernel."""
import os
import re

import numpy as np
import pytest

import comfec as F
import oracle as O
from conftest import GOLDEN, ROOT

SIONNA = dict(np.load(os.path.join(GOLDEN, "sionna_ldpc.npz"), allow_pickle=False))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "comfec_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = F.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    # and the Python view binds each of them
    bound = {name for name, _, _ in F.SIGNATURES}
    assert set(syms) <= bound


@pytest.mark.parametrize("k,n", [(200, 220), (1347, 1554), (8448, 25344), (100, 300)] +
                         [tuple(int(x) for x in key.split(".")[1][1:].split("_n"))
                          for key in sorted(SIONNA) if key.endswith(".u")][::4])
def test_encoder_matches_sionna_and_oracle(k, n):
    code = F.NrLdpc(k, n)
    ref = O.Code.nr(k, n)
    assert (code.base_graph(), code.lifting_size()) == (ref.bg, ref.z)
    assert np.array_equal(code.bit_map(), ref.bit_map())
    key = f"ldpc.k{k}_n{n}"
    if key + ".u" in SIONNA:
        u, c = SIONNA[key + ".u"], SIONNA[key + ".c"]
        for i in range(len(u)):
            assert np.array_equal(code.encode_block(u[i]), c[i])
    info = np.random.default_rng(k).integers(0, 2, k).astype(np.uint8)
    assert np.array_equal(code.encode_block(info), ref.nr_encode(info))


@pytest.mark.parametrize("k,e,rv,qm", [(100, 300, 0, 1), (500, 1200, 2, 4), (1000, 1500, 3, 6), (60, 1000, 1, 2),
                                       (8448, 25344, 0, 4), (3840, 11520, 1, 2)])
def test_rate_matching_matches_oracle(k, e, rv, qm):
    code = F.NrLdpc(k, e, None, rv, qm)
    ref = O.Code.nr(k, e, 0, rv, qm)
    assert np.array_equal(code.bit_map(), ref.bit_map())


def test_lifted_codes_match_oracle():
    for bg, z in [(2, 52), (1, 384), (2, 16), (1, 15)]:
        code = F.NrLdpc.lifted(bg, z)
        ref = O.Code.lifted(bg, z)
        assert (code.dims.n, code.dims.k, code.dims.e, code.dims.edges) == (ref.n, ref.k, ref.n - 2 * z, ref.edges)
        u = np.random.default_rng(z).integers(0, 2, ref.k).astype(np.uint8)
        assert np.array_equal(code.encode_block(u), ref.encode(u)[2 * z:])


def test_errors_carry_the_reference_kinds_and_messages():
    with pytest.raises(F.ComfecError) as e:
        F.NrLdpc(9000, 10000)
    assert e.value.kind == "Config" and "exceed one code block" in str(e.value)
    with pytest.raises(F.ComfecError) as e:
        F.NrLdpc(100, 301, None, 0, 2)
    assert e.value.kind == "Config" and "Qm | E" in str(e.value)
    with pytest.raises(F.ComfecError) as e:
        F.NrLdpc.lifted(3, 52)
    assert e.value.kind == "Config"
    code = F.NrLdpc(100, 300)
    with pytest.raises(F.ComfecError) as e:
        code.encode_block(np.zeros(99, np.uint8))
    assert e.value.kind == "Shape" and str(e.value) == "inconsistent shapes: 99 bits for a block of 100"
    with pytest.raises(F.ComfecError) as e:
        code.decode_block(np.zeros(10))
    assert e.value.kind == "Shape" and str(e.value) == "inconsistent shapes: 10 LLRs for a block of 300"
    with pytest.raises(F.ComfecError) as e:
        code.set_decoder("belief", 0.5)
    assert e.value.kind == "Config"
    assert F.rule_from_name("MIN_SUM", 0.3) == (F.NORMALIZED_MIN_SUM, 1.0)
