"""Copy the reference-held data fixtures of the NR LDPC path into tests/golden/.

Run once in a container that has the reference checked out at /root/reference; the
outputs are committed, and nothing at test time reads the reference.

- sionna_ldpc.npz: the ``ldpc.*`` entries of tests/oracle/sionna_vectors.npz (Sionna,
  Apache-2.0, commit recorded in sionna_ldpc.json): NR LDPC encoder + rate matching
  codewords for 28 (K', E), two messages each (u: info bits, c: transmitted bits).
- aff3ct_*.json: the LDPC entries of validation/aff3ct/ (AFF3CT v4.7.1 FER/BER curves).
- *.alist: the PEG matrices those AFF3CT runs decoded.
"""
import json
import os
import shutil

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    v = np.load(os.path.join(REF, "tests/oracle/sionna_vectors.npz"), allow_pickle=False)
    ldpc = {k: v[k] for k in v.files if k.startswith("ldpc.")}
    np.savez_compressed(os.path.join(OUT, "sionna_ldpc.npz"), **ldpc)
    meta = json.load(open(os.path.join(REF, "tests/oracle/sionna_vectors.json")))
    keep = {k: meta[k] for k in meta if not isinstance(meta[k], dict) or k in ("source",)}
    json.dump(keep, open(os.path.join(OUT, "sionna_ldpc.json"), "w"), indent=1, sort_keys=True)
    for name in sorted(os.listdir(os.path.join(REF, "validation/aff3ct"))):
        if name.startswith("ldpc_") and name.endswith(".json"):
            shutil.copy(os.path.join(REF, "validation/aff3ct", name), os.path.join(OUT, "aff3ct_" + name))
    for name in ("peg_n576_k288.alist", "peg_n648_k540.alist"):
        shutil.copy(os.path.join(REF, "validation/aff3ct/matrices", name), os.path.join(OUT, name))


if __name__ == "__main__":
    main()
