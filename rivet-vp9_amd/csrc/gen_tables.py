#!/usr/bin/env python3
"""Regenerates vp9_tables.h: the numeric tables of the VP9 bitstream
specification (probabilities, scans, quantiser lookups, filter taps, block
lookups) as C arrays.

The tables are data fixed by the specification.  They are read, as text, from
the reference crate's transcription (src/tables.rs of safewords/rivet-vp9,
itself generated from the specification PDF) so that both sides hold the same
numbers; nothing else of that file is used.  Usage:

    python3 gen_tables.py /root/reference/src/tables.rs > vp9_tables.h

The output records the SHA-256 of the text it was made from and a checksum of
every table, which tests/test_tables.py re-checks.
"""
import hashlib
import re
import sys

RUST_TO_C = {"u8": "uint8_t", "i8": "int8_t", "u16": "uint16_t", "i32": "int32_t"}
ALIASES = {"CoefProbs": "[[[[[[u8; 3]; 6]; 6]; 2]; 2]; 4]"}


def parse_type(t):
    """'[[i32; 256]; 3]' -> ('i32', [3, 256])."""
    t = ALIASES.get(t.strip(), t.strip())
    dims = []
    while t.startswith("["):
        inner = t[1:-1]
        # split at the last ';' at depth 0
        depth = 0
        cut = None
        for i, ch in enumerate(inner):
            if ch == "[":
                depth += 1
            elif ch == "]":
                depth -= 1
            elif ch == ";" and depth == 0:
                cut = i
        elem, n = inner[:cut], inner[cut + 1:]
        dims.append(int(n.strip()))
        t = elem.strip()
    return t, dims


def main(path):
    text = open(path, encoding="utf-8").read()
    digest = hashlib.sha256(text.encode()).hexdigest()
    decl = re.compile(r"pub\(crate\) static (\w+): ([^=]+)= (\[.*?\]);\s*$", re.M)
    out = []
    out.append("/* Numeric tables of the VP9 bitstream specification (v0.7, sections 6, 8,")
    out.append(" * 10), as C arrays.  Generated by gen_tables.py from the reference crate's")
    out.append(" * transcription src/tables.rs; do not edit by hand.")
    out.append(f" * source sha256: {digest} */")
    out.append("#ifndef VP9R_TABLES_H")
    out.append("#define VP9R_TABLES_H")
    out.append("#include <stdint.h>")
    out.append("")
    sums = []
    for m in decl.finditer(text):
        name, rtype, body = m.group(1), m.group(2), m.group(3)
        elem, dims = parse_type(rtype)
        vals = [int(v) for v in re.findall(r"-?\d+", body)]
        total = 1
        for d in dims:
            total *= d
        if len(vals) != total:
            raise SystemExit(f"{name}: {len(vals)} values for {dims}")
        ctype = RUST_TO_C[elem]
        shape = "".join(f"[{d}]" for d in dims)
        chk = sum((i + 1) * v for i, v in enumerate(vals)) & 0xFFFFFFFF
        sums.append((name, chk))
        out.append(f"/* {name}: {rtype.strip()} (weighted checksum {chk}) */")
        out.append(f"static const {ctype} VP9_{name}{shape} = {{")
        for i in range(0, len(vals), 24):
            out.append("  " + ", ".join(str(v) for v in vals[i:i + 24]) + ",")
        out.append("};")
        out.append("")
    out.append("#endif")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/tables.rs")
