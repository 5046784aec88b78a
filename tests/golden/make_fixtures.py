#!/usr/bin/env python3
"""Turn the reference's consensus-spec vectors into compact binary fixtures.

Inputs (read as data only, from the reference checkout):
  tests/vectors/*/data.yaml   -- consensus-spec compute/recover vectors
  src/setup_data.cpp          -- the mainnet G1 monomial trusted setup bytes

Outputs (committed):
  tests/golden/compute_vectors.bin  -- blob(131072) [+ proofs(6144)] records
  tests/golden/vectors.json         -- index: offsets, sha256 of cells, recover cases
  vkzg_amd/data/g1_monomial_mainnet.bin -- 4096 x 48-byte compressed G1 points

Cells are not stored: the oracle recomputes them from the blob and the tests
compare the sha256 recorded here.  Every valid recover vector's input cells
and expected output are cells of one of the compute vectors; this script
checks that and records which one.

Usage: python3 tests/golden/make_fixtures.py [/root/reference]
"""
import hashlib
import json
import os
import re
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BLOB = 131072
PROOFS = 128 * 48
CELL = 2048
SETUP_SHA = "08797579f6cfd5788eddc1a215d64dcfabd04acbcaf2953fb2c1afb830f43315"


def hexbytes(s):
    assert s.startswith("0x")
    return bytes.fromhex(s[2:]) if len(s) % 2 == 0 else None


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    vdir = os.path.join(ref, "tests", "vectors")
    index = {"compute": [], "recover": [], "provenance": {
        "source": "consensus-spec vectors shipped in the reference under tests/vectors",
        "setup_sha256": SETUP_SHA}}
    payload = bytearray()
    cells_by_name = {}
    proofs_by_name = {}

    for i in range(4):
        name = f"compute_cells_and_kzg_proofs_case_invalid_blob_{i}"
        d = yaml.safe_load(open(os.path.join(vdir, name, "data.yaml")))
        assert d["output"] is None
        raw = d["input"]["blob"]
        b = hexbytes(raw)
        rec = {"name": name, "valid": False, "hex_chars": len(raw) - 2}
        if b is not None and len(b) == BLOB:
            rec["offset"] = len(payload)
            rec["blob_len"] = BLOB
            payload += b
        else:
            # Wrong-length blobs are rejected by the caller before the library
            # is reached; keep the length (in bytes, possibly fractional) only.
            rec["blob_len"] = (len(raw) - 2) / 2
        index["compute"].append(rec)

    for i in range(7):
        name = f"compute_cells_and_kzg_proofs_case_valid_{i}"
        d = yaml.safe_load(open(os.path.join(vdir, name, "data.yaml")))
        b = hexbytes(d["input"]["blob"])
        assert len(b) == BLOB
        cells = b"".join(hexbytes(c) for c in d["output"][0])
        proofs = b"".join(hexbytes(p) for p in d["output"][1])
        assert len(cells) == 128 * CELL and len(proofs) == PROOFS
        rec = {"name": name, "valid": True, "offset": len(payload), "blob_len": BLOB,
               "cells_sha256": hashlib.sha256(cells).hexdigest()}
        payload += b
        rec["proofs_offset"] = len(payload)
        payload += proofs
        index["compute"].append(rec)
        cells_by_name[name] = cells
        proofs_by_name[name] = proofs

    for name in sorted(os.listdir(vdir)):
        if not name.startswith("recover_cells_and_kzg_proofs_case_"):
            continue
        d = yaml.safe_load(open(os.path.join(vdir, name, "data.yaml")))
        idx = [int(x) for x in d["input"]["cell_indices"]]
        cin = [hexbytes(c) for c in d["input"]["cells"]]
        src = None
        for cname, cells in cells_by_name.items():
            if all(cells[k * CELL:(k + 1) * CELL] == c for k, c in zip(idx, cin)):
                src = cname
                break
        rec = {"name": name, "valid": d["output"] is not None, "cell_indices": idx,
               "source": src}
        if d["output"] is not None:
            out = b"".join(hexbytes(c) for c in d["output"][0])
            outp = b"".join(hexbytes(p) for p in d["output"][1])
            assert src is not None and out == cells_by_name[src], name
            rec["cells_sha256"] = hashlib.sha256(out).hexdigest()
            rec["proofs_equal_source"] = outp == proofs_by_name[src]
        elif idx:
            assert src is not None, name
        index["recover"].append(rec)

    with open(os.path.join(HERE, "compute_vectors.bin"), "wb") as f:
        f.write(payload)
    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(index, f, indent=1)

    # Mainnet setup: the byte array in setup_data.cpp, as data.
    text = open(os.path.join(ref, "src", "setup_data.cpp")).read()
    body = text[text.index("{", text.index("kEmbeddedSetupG1Monomial")) + 1:]
    body = body[:body.index("}")]
    setup = bytes(int(h, 16) for h in re.findall(r"0x([0-9a-fA-F]{2})", body))
    assert len(setup) == 196608, len(setup)
    assert hashlib.sha256(setup).hexdigest() == SETUP_SHA
    os.makedirs(os.path.join(REPO, "vkzg_amd", "data"), exist_ok=True)
    with open(os.path.join(REPO, "vkzg_amd", "data", "g1_monomial_mainnet.bin"), "wb") as f:
        f.write(setup)
    print("wrote", len(payload), "fixture bytes;", len(index["recover"]), "recover cases")


if __name__ == "__main__":
    main()
