"""Named Gaussian basis sets, as data (src/Data.jl:5-1184 of the reference).

``basis_data[name][atno]`` is a list of ``(sym, [(expn, coef), ...])`` shells,
the same shape ``shells()`` iterates in src/Basis.jl:100-120.  The numbers live
in data/basis_library.json (regenerated by data/make_basis_library.py).
"""
import json
import os

_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis_library.json")
with open(_path) as _f:
    _raw = json.load(_f)

basis_data = {name: {int(z): [(sym, [tuple(p) for p in prims]) for sym, prims in shells]
                     for z, shells in per_atom.items()}
              for name, per_atom in _raw.items()}
