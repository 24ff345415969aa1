"""The DVB-S2 LDPC code of the short FECFRAME, rate 1/2 (N = 16200, K = 7200), built from
the parity accumulator addresses of ETSI EN 302 307-1 Annex C (tests/golden/dvbs2_n16200_k7200.json)
the way crates/com-systems/src/dvbs2/fec.rs:213-243 builds its rows: information bit
360 g + l enters check (x + l q) mod (N - K) for every address x of group g, q = (N - K) / 360,
and check r also holds the parity bits K + r - 1 (r > 0) and K + r (the accumulator)."""
import json
import os

from conftest import GOLDEN


def dvbs2_short_r1_2_rows():
    fx = json.load(open(os.path.join(GOLDEN, "dvbs2_n16200_k7200.json")))
    n, k, group, table = fx["n"], fx["k"], fx["group"], fx["table"]
    m = n - k
    q = m // group
    rows = [[] for _ in range(m)]
    bit = i = 0
    while i < len(table):
        count = table[i]
        addresses = table[i + 1:i + 1 + count]
        for l in range(group):
            for x in addresses:
                rows[(x + l * q) % m].append(bit)
            bit += 1
        i += 1 + count
    assert bit == k
    for r in range(m):
        if r > 0:
            rows[r].append(k + r - 1)
        rows[r].append(k + r)
    return n, k, rows
