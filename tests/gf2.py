"""GF(2) helpers of the tests: the non-pivot columns of a parity-check matrix's reduced form
with the pivots taken from the last column (the order of crates/com-math/src/gf2.rs:124-150;
a reduced form is unique for its pivot order)."""
import numpy as np


def nonpivot_columns(n, rows):
    h = np.zeros((len(rows), n), np.uint8)
    for r, lst in enumerate(rows):
        for v in lst:
            h[r, v] ^= 1
    rank, pivots = 0, set()
    for c in range(n - 1, -1, -1):
        if rank == len(rows):
            break
        hit = np.nonzero(h[rank:, c])[0]
        if len(hit) == 0:
            continue
        p = rank + hit[0]
        h[[rank, p]] = h[[p, rank]]
        for r in np.nonzero(h[:, c])[0]:
            if r != rank:
                h[r] ^= h[rank]
        pivots.add(c)
        rank += 1
    return [c for c in range(n) if c not in pivots]
