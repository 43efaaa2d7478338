"""Extract the baked hex P3/P4 DFR operator constants from the reference's GiMMiK kernels.

The reference's GPU operator products (src/gimmik_gpu.cu) are generated code: each
``gimmik_mm_gpu_<id>_b0`` kernel hard-codes the nonzero entries of one constant operator
A as ``dotp = a*b[i + k*ldb] + ...; c[i + r*ldc] = dotp;``.  The kernel id is
``hash_str("opp<name>_P<p>_<D>D_e<etype>")`` (src/elements.cpp:750-773) with the djb2
string hash of src/funcs.cpp:1536-1545.  Only the hexahedron (etype 4) entries for P3
and P4 were generated with the current element-type numbering, so only those are
usable as golden operator values.

This script reads the reference source as text and writes the operator entries
(row, column, value) to ``gimmik_hex_p3p4.json`` next to it.  It needs the reference
checkout (default ``/root/reference``); the JSON it writes is the committed fixture.

    python tests/golden/extract_gimmik.py [/path/to/reference]
"""

import json
import os
import re
import sys

OPS = {  # name -> (rows, cols) as functions of nS (spts) and nF (fpts), D = 3
    "oppE": lambda nS, nF: (nF, nS),
    "oppD": lambda nS, nF: (3 * nS, nS),
    "oppD_fpts": lambda nS, nF: (3 * nS, nF),
    "oppDiv": lambda nS, nF: (nS, 3 * nS),
    "oppDiv_fpts": lambda nS, nF: (nS, nF),
}


def djb2(s):
    h = 5381
    for ch in s.encode():
        h = (h * 33 + ch) & 0xFFFFFFFFFFFFFFFF
    return h


def parse_kernel(src, kid):
    m = re.search(r"gimmik_mm_gpu_%d_b0\(.*?\{(.*?)\n\}" % kid, src, re.S)
    if m is None:
        raise KeyError(kid)
    body = m.group(1)
    entries = []
    term = re.compile(r"(-?[0-9.]+(?:[eE][-+]?[0-9]+)?)\*b\[i \+ (\d+)\*ldb\]")
    for stmt in re.finditer(r"dotp = (.*?);\s*c\[i \+ (\d+)\*ldc\] = dotp;", body, re.S):
        expr, row = stmt.group(1), int(stmt.group(2))
        for t in term.finditer(expr):
            coef = float(t.group(1))
            entries.append([row, int(t.group(2)), coef])
    zero_rows = [int(r) for r in re.findall(r"c\[i \+ (\d+)\*ldc\] = 0(?:\.0)?;", body)]
    return entries, zero_rows


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "gimmik_gpu.cu")) as f:
        src = f.read()
    out = {}
    for p in (3, 4):
        nS, nF = (p + 1) ** 3, 6 * (p + 1) ** 2
        for name, shape in OPS.items():
            key = "%s_P%d_3D_e4" % (name, p)
            kid = djb2(key)
            entries, zero_rows = parse_kernel(src, kid)
            out[key] = {"id": str(kid), "shape": list(shape(nS, nF)), "entries": entries,
                        "zero_rows": zero_rows}
            print("%-22s id=%d nnz=%d" % (key, kid, len(entries)))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gimmik_hex_p3p4.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", dst)


if __name__ == "__main__":
    main()
