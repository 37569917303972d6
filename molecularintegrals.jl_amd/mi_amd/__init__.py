"""mi_amd -- MI355X-native Head-Gordon/Pople ERI engine, Python host side.

Mirrors MolecularIntegrals.jl's interface for its HGP two-electron path:
the basis model (``Basis``, ``build_basis``, ``pgbf``, ``cgbf``, ``addbf``),
the index algebra (``iindex``, ``iiterator``) and the drivers
``all_twoe_ints_chrr`` / ``all_twoe_ints``, which run on the GPU through
libmi_hgp.so (C ABI in include/mi_hgp.h).
"""
from .basis import (Atom, Basis, CGBF, FlatBasis, PGBF, Shell, addbf, atom, basis_from_shells, build_basis, cgbf,
                    flat_from_shells, flatten, m2ao, nao, overlap, pgbf, shell_functions, shell_indices,
                    sub_basis)
from .indexing import chrr_index, iindex, iiterator, n_unique, ORDER_CHRR, ORDER_IINDEX
from .eri import Context, Plan, all_twoe_ints, all_twoe_ints_chrr, default_context

__all__ = ["Atom", "Basis", "CGBF", "FlatBasis", "PGBF", "Shell", "addbf", "atom", "basis_from_shells",
           "build_basis", "cgbf", "flat_from_shells", "flatten", "m2ao", "nao", "overlap", "pgbf", "shell_functions", "shell_indices", "sub_basis",
           "chrr_index", "iindex", "iiterator", "n_unique", "ORDER_CHRR", "ORDER_IINDEX", "Context", "Plan",
           "all_twoe_ints", "all_twoe_ints_chrr", "default_context"]
