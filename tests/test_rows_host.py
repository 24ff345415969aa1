"""Codes given by the rows of their parity-check matrix (LdpcCode::from_rows,
crates/com-fec/src/ldpc.rs:216-231), host side of the C ABI (no GPU needed).

- K and the information positions: the (7,4) Hamming code of the reference's test
  (ldpc.rs:1066-1072) has K = 4; on the PEG matrices of the AFF3CT fixtures K = N - rank H
  and the information positions are the columns that are not pivots when the reduced form
  takes its pivots from the last column (the reduction of crates/com-math/src/gf2.rs:124-150;
  a reduced form is unique for a pivot order, restated here by plain elimination);
- the encoder: H c = 0 on the oracle's copy of the matrix for random information bits,
  which sit unchanged at the information positions;
- errors: a variable out of range (the reference's second assertion), rows too long for
  the device kernel.
"""
import os

import numpy as np
import pytest

import comfec as F
import oracle as O
from alist import read_alist
from conftest import GOLDEN
from gf2 import nonpivot_columns


def test_hamming_code_of_the_reference():
    code = F.NrLdpc.from_rows(7, [[0, 1, 2, 4], [0, 1, 3, 5], [0, 2, 3, 6]])
    assert code.k == 4 and code.n == 7 and code.dims.base_graph == 0
    with pytest.raises(F.ComfecError) as e:
        F.NrLdpc.from_rows(3, [[0, 5]])
    assert e.value.kind == "Config" and "missing variable" in str(e.value)


@pytest.mark.parametrize("matrix", ["peg_n576_k288.alist", "peg_n648_k540.alist"])
def test_peg_codes_information_positions_and_encoder(matrix):
    n, rows = read_alist(os.path.join(GOLDEN, matrix))
    code = F.NrLdpc.from_rows(n, rows)
    info = nonpivot_columns(n, rows)
    assert code.k == len(info) and code.n == n
    ref = O.Code.from_rows(n, rows)
    rng = np.random.default_rng(3)
    for _ in range(4):
        u = rng.integers(0, 2, code.k).astype(np.uint8)
        cw = code.encode_block(u)
        assert ref.is_codeword(cw)
        assert np.array_equal(cw[info], u)


def test_rows_too_long_for_the_device():
    with pytest.raises(F.ComfecError) as e:
        F.NrLdpc.from_rows(40, [list(range(20)), list(range(20, 40))])
    assert e.value.kind == "Config"


def test_dvbs2_short_code_structure_and_encoder():
    """The DVB-S2 short rate-1/2 code from the standard's table: the degree distribution of
    the standard (1800 information bits of degree 8, 5400 of degree 3, the accumulator's
    parity bits of degree 2 but the last of degree 1; checks of 4 to 7 bits), K = 7200 with
    the information bits first (the accumulator is lower triangular, so the parity
    columns are the pivots), and codewords of the host encoder satisfy H c = 0."""
    from dvbs2 import dvbs2_short_r1_2_rows
    n, k, rows = dvbs2_short_r1_2_rows()
    deg = np.zeros(n, int)
    for r in rows:
        deg[r] += 1
    assert (np.sum(deg[:k] == 8), np.sum(deg[:k] == 3)) == (1800, 5400)
    assert np.all(deg[k:n - 1] == 2) and deg[n - 1] == 1
    assert {len(r) for r in rows} == {4, 5, 6, 7} and sum(len(r) for r in rows) == 48599
    code = F.NrLdpc.from_rows(n, rows)
    assert code.k == k and code.n == n
    ref = O.Code.from_rows(n, rows)
    u = np.random.default_rng(5).integers(0, 2, k).astype(np.uint8)
    cw = code.encode_block(u)
    assert ref.is_codeword(cw) and np.array_equal(cw[:k], u)
