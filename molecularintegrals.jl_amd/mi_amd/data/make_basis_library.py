"""Regenerates basis_library.json (the named Gaussian basis sets: sto3g, 6-31G,
6-31G**, cc-pvdz) from the reference's data tables, src/Data.jl:5-1184.

Run in a checkout that has the reference available:
    python make_basis_library.py /path/to/MolecularIntegrals.jl/src/Data.jl
The JSON holds only numbers: {name: {atno: [[sym, [[expn, coef], ...]], ...]}}.
"""
import ast
import json
import os
import re
import sys

NAMES = {"sto3g": "sto3g", "g631ss": "6-31G**", "g631": "6-31G", "ccpvdz": "cc-pvdz"}


def parse(text):
    out = {}
    for var, name in NAMES.items():
        m = re.search(r"^%s = \[" % var, text, re.M)
        start = m.end() - 1
        depth, i = 0, start
        while True:
            ch = text[i]
            if ch == "[":
                depth += 1
            elif ch == "]":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        body = re.sub(r"#[^\n]*", "", text[start:i + 1])
        arr = ast.literal_eval(body)
        out[name] = {str(z + 1): [[sym, [list(p) for p in prims]] for sym, prims in atomdata]
                     for z, atomdata in enumerate(arr)}
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/Data.jl"
    with open(src) as f:
        data = parse(f.read())
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_library.json")
    with open(dst, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("wrote", dst, {k: len(v) for k, v in data.items()})
