#!/usr/bin/env python3
"""Generates provider/types.inc (C++) and provider/types.json (Python hosts): the
payload type table the provider modules check descriptors against, from a
KSpaceJet type registry (types/registry.json).

A type identity is SHA-256 over b"kspacejet.type-identity\\0" followed by the
canonical JSON (sorted keys, no spaces) of the flattened structural descriptor
{type_ref, payload_kind, element_type, rank, dimensions, layout, strides,
explicit_byte_strides, allowed_memory_domains, min_alignment_bytes, mutability}
(the rule of the reference's tools/type_registry/generate.py:225-227,497-509).

    python kspacejet_amd/csrc/provider/gen_types.py /path/to/types/registry.json
"""
import hashlib
import json
import os
import sys

TYPES = ["ksj.kspace-frame", "ksj.noncartesian-kspace-frame", "ksj.trajectory-frame", "ksj.coil-image-frame",
         "ksj.image-frame", "ksj.complex-image-frame", "ksj.sensitivity-map", "ksj.noise-model", "ksj.phase-model",
         "ksj.coil-compression-basis", "ksj.noise-calibration-frame", "ksj.phase-reference-frame"]
ELEMENT = {"float32": "KSJ_ELEMENT_TYPE_FLOAT32", "complex_float32": "KSJ_ELEMENT_TYPE_COMPLEX_FLOAT32"}
KIND = {"buffer_handle": "KSJ_PAYLOAD_KIND_BUFFER_HANDLE"}
LAYOUT = {"channel_major_contiguous": "KSJ_TYPE_LAYOUT_CHANNEL_MAJOR_CONTIGUOUS",
          "row_major_contiguous": "KSJ_TYPE_LAYOUT_ROW_MAJOR_CONTIGUOUS"}
STRIDES = {"canonical": "KSJ_TYPE_STRIDES_CANONICAL"}
DOMAIN = {"host_normal": "KSJ_PROVIDER_MEMORY_HOST_PAGEABLE"}
MUT = {"immutable_after_publish": "KSJ_PAYLOAD_MUTABILITY_IMMUTABLE_AFTER_PUBLISH"}


def canonical(v):
    return json.dumps(v, ensure_ascii=False, sort_keys=True, separators=(",", ":"), allow_nan=False)


def digest(t):
    p, m = t["identity"]["payload"], t["identity"]["memory"]
    flat = {"type_ref": t["type_ref"], "payload_kind": p["kind"], "element_type": p["element_type"], "rank": p["rank"],
            "dimensions": p["dimensions"], "layout": p["layout"], "strides": p["strides"],
            "explicit_byte_strides": p["explicit_byte_strides"], "allowed_memory_domains": m["allowed_domains"],
            "min_alignment_bytes": m["min_alignment_bytes"], "mutability": t["identity"]["mutability"]}
    return hashlib.sha256(b"kspacejet.type-identity\0" + canonical(flat).encode()).digest()


def main(registry):
    reg = json.load(open(registry))
    by_ref = {t["type_ref"]: t for t in reg["types"]}
    table = []
    out = ["// Generated by gen_types.py from the KSpaceJet type registry.  Do not edit.",
           "// One aggregate initializer per type: {type_ref, identity digest, kind, element, rank,",
           "// dimension names, layout | strides, memory domains, alignment, mutability}."]
    for ref in TYPES:
        t = by_ref[ref]
        p, m = t["identity"]["payload"], t["identity"]["memory"]
        assert p["strides"] == "canonical" and not p["explicit_byte_strides"], ref  # stride_bytes stay zero
        d = digest(t)
        dims = ", ".join(f'"{x}"' for x in p["dimensions"])
        table.append({"type_ref": ref, "digest": d.hex(), "kind": p["kind"], "element_type": p["element_type"],
                      "rank": p["rank"], "dimensions": p["dimensions"], "layout": p["layout"],
                      "strides": p["strides"], "allowed_domains": m["allowed_domains"],
                      "min_alignment_bytes": m["min_alignment_bytes"], "mutability": t["identity"]["mutability"]})
        out.append("{" + ", ".join([
            f'"{ref}"', "{" + ", ".join(f"0x{b:02x}" for b in d) + "}", KIND[p["kind"]], ELEMENT[p["element_type"]],
            str(p["rank"]), "{" + dims + "}", f'{LAYOUT[p["layout"]]} | {STRIDES[p["strides"]]}',
            " | ".join(DOMAIN[x] for x in m["allowed_domains"]), str(m["min_alignment_bytes"]),
            MUT[t["identity"]["mutability"]]]) + "},")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "types.inc")
    open(path, "w").write("\n".join(out) + "\n")
    jpath = os.path.join(os.path.dirname(path), "types.json")
    json.dump(table, open(jpath, "w"), indent=1)
    print("wrote", path, "and", jpath)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/types/registry.json")
