"""Copy the reference-held data fixtures of the NR LDPC path into tests/golden/.

Run once in a container that has the reference checked out at /root/reference; the
outputs are committed, and nothing at test time reads the reference.

- sionna_ldpc.npz: the ``ldpc.*`` entries of tests/oracle/sionna_vectors.npz (Sionna,
  Apache-2.0, commit recorded in sionna_ldpc.json): NR LDPC encoder + rate matching
  codewords for 28 (K', E), two messages each (u: info bits, c: transmitted bits).
- aff3ct_*.json: the LDPC entries of validation/aff3ct/ (AFF3CT v4.7.1 FER/BER curves).
- *.alist: the PEG matrices those AFF3CT runs decoded.
- dvbs2_n16200_k7200.json: the parity accumulator address table of the DVB-S2 LDPC code
  of the short FECFRAME, rate 1/2 (ETSI EN 302 307-1 V1.4.1, Annex C), as the reference
  holds it (crates/com-systems/src/dvbs2/tables.rs, N16200_P9000): the code of the AFF3CT
  run ldpc_dvbs2_short_r1_2_layered_spa_i50.
- gray_mapping_qam_{4,16,64}.json: the Gray-mapped square QAM points of OptiCommPy 0.11.0
  (tests/oracle/gray_mapping_qam_*.json, used by tests/test_oracle.py:97-105): point i of
  the unnormalised constellation is the symbol of label i.
"""
import json
import os
import re
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
    for m in (4, 16, 64):
        name = f"gray_mapping_qam_{m}.json"
        shutil.copy(os.path.join(REF, "tests/oracle", name), os.path.join(OUT, name))
    text = open(os.path.join(REF, "crates/com-systems/src/dvbs2/tables.rs")).read()
    body = re.search(r"const N16200_P9000: &\[u16\] = &\[(.*?)\];", text, re.S).group(1)
    table = [int(x) for x in body.split(",") if x.strip()]
    json.dump({"n": 16200, "k": 7200, "group": 360,
               "source": "ETSI EN 302 307-1 V1.4.1 Annex C, short FECFRAME, rate 1/2 "
                         "(crates/com-systems/src/dvbs2/tables.rs N16200_P9000)",
               "layout": "per group of 360 information bits: the count of addresses, then the addresses",
               "table": table}, open(os.path.join(OUT, "dvbs2_n16200_k7200.json"), "w"))


if __name__ == "__main__":
    main()
