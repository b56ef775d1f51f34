#!/usr/bin/env python3
"""Writes type_registry_reference.json: for each payload type the provider
modules use, its identity object from the reference's types/registry.json and
the 32-byte identity digest the reference SDK's generated type_registry.h holds
for it.  The fixture lets tests recompute each digest from the identity with
this repository's rule and compare it with the reference's own bytes.

    python tests/golden/gen_type_registry_fixture.py /root/reference
"""
import json
import os
import re
import sys

TYPES = ["ksj.kspace-frame", "ksj.noncartesian-kspace-frame", "ksj.trajectory-frame", "ksj.coil-image-frame",
         "ksj.image-frame", "ksj.complex-image-frame", "ksj.sensitivity-map", "ksj.noise-model", "ksj.phase-model",
         "ksj.coil-compression-basis", "ksj.noise-calibration-frame", "ksj.phase-reference-frame"]


def main(root):
    reg = {t["type_ref"]: t for t in json.load(open(os.path.join(root, "types/registry.json")))["types"]}
    header = open(os.path.join(root, "libs/recon/kspacejet-provider-sdk/include/kspacejet/provider/type_registry.h")).read()
    out = []
    for ref in TYPES:
        stem = ref.replace("ksj.", "").replace("-", "_")
        i = header.index(f"ksj_type_registry_{stem}_identity_digest =")
        digest = bytes(int(x, 16) for x in re.findall(r"UINT8_C\((0x[0-9a-f]+)\)", header[i:i + 3000])[:32])
        out.append({"type_ref": ref, "identity": reg[ref]["identity"], "header_digest": digest.hex()})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "type_registry_reference.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
