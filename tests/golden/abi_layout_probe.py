#!/usr/bin/env python3
"""Struct layout (sizeof, alignof, offsetof of every field) of the Provider ABI.

`probe(header)` compiles a small C program against a header that declares the
ABI and returns {struct: {"size", "align", "fields": {name: offset}}}.  The
struct and field list is read from include/ksj_provider_abi.h.

Run as a script, in a tree that holds the reference SDK, it writes the layout
of the reference's provider.h to abi_layout_reference.json beside it:

    python tests/golden/abi_layout_probe.py /root/reference
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
OURS = os.path.join(HERE, "..", "..", "include", "ksj_provider_abi.h")
REF_REL = "libs/recon/kspacejet-provider-sdk/include/kspacejet/provider/provider.h"


def structs(header_text: str):
    """[(name, [field, ...])] of every struct body in the header."""
    out = []
    for m in re.finditer(r"(?:typedef\s+)?struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)?\s*;", header_text, re.S):
        name, body = m.group(1), m.group(2)
        fields = []
        for line in body.split(";"):
            line = re.sub(r"/\*.*?\*/", "", line, flags=re.S).strip()
            if not line:
                continue
            fm = re.search(r"\(\s*KSJ_PROVIDER_CALL\s*\*\s*(\w+)\s*\)", line) or re.search(r"(\w+)\s*(\[[^\]]*\])?\s*$", line)
            fields.append(fm.group(1))
        out.append((name, fields))
    return out


def probe(header_path: str, include_dirs=()):
    spec = structs(open(OURS).read())
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{header_path}"', "int main(void) {",
             '  printf("{");']
    for i, (name, fields) in enumerate(spec):
        sep = "," if i else ""
        lines.append(f'  printf("{sep}\\"{name}\\":{{\\"size\\":%zu,\\"align\\":%zu,\\"fields\\":{{", '
                     f"sizeof(struct {name}), _Alignof(struct {name}));")
        for j, f in enumerate(fields):
            fsep = "," if j else ""
            lines.append(f'  printf("{fsep}\\"{f}\\":%zu", offsetof(struct {name}, {f}));')
        lines.append('  printf("}}");')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "p.c"), os.path.join(d, "p")
        open(src, "w").write("\n".join(lines))
        cmd = ["gcc", "-std=c11", "-o", exe, src] + [f"-I{x}" for x in include_dirs]
        subprocess.run(cmd, check=True, capture_output=True, text=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    ref = os.path.join(root, REF_REL)
    layout = probe(ref)
    path = os.path.join(HERE, "abi_layout_reference.json")
    json.dump(layout, open(path, "w"), indent=0, sort_keys=True)
    print("wrote", path, len(layout), "structs")
