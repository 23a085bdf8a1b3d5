#!/usr/bin/env python3
"""Extract the in-tree Poseidon2 width-16 constants the reference ships
(integration_test/zkc/riscv/zkc/lib/poseidon2/constants.zkc: round keys
:34-62, internal diagonal :77-80) into tests/golden/poseidon2_w16_zkc.json.

These are data (27 x 16 round keys and 16 diagonal entries, canonical
KoalaBear values) -- the only absolute anchor the reference tree holds for
gnark-crypto's Poseidon2 constants.  Run in a container that has the
reference mounted at /root/reference.
"""
import json
import os
import re

SRC = "/root/reference/integration_test/zkc/riscv/zkc/lib/poseidon2/constants.zkc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "poseidon2_w16_zkc.json")


def main():
    text = open(SRC).read()
    rk = text.split("static round_keys")[1].split("}")[0]
    dg = text.split("static internal_diag")[1].split("}")[0]
    keys = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]{8}", rk)]
    diag = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]{8}", dg)]
    assert len(keys) == 27 * 16 and len(diag) == 16
    with open(OUT, "w") as f:
        json.dump({"source": "integration_test/zkc/riscv/zkc/lib/poseidon2/constants.zkc",
                   "round_keys_27x16": keys, "internal_diag": diag}, f)
    print(OUT)


if __name__ == "__main__":
    main()
