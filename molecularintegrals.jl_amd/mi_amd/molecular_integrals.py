"""Drop-in for the reference's Python entry point (python/molecular_integrals.py
of rpmuller/MolecularIntegrals.jl) on its two-electron path, routed to the
GPU HGP engine instead of Julia.

The reference wrapper drives Julia through juliacall; ``all_twoe_ints(basis,
method)`` (python/molecular_integrals.py:409-440) passes ``coulomb`` for
'default' and, despite the name, also for 'hgp' (:427-430), so its Python API
never reaches the HGP code.  Here both names run the HGP recurrences on the
GPU (C ABI ``mi_eri_unique``, iindex order, src/ERI.jl:104-112) and return the
same values, the normalised contracted coulomb() of every unique quartet.

Names and argument meaning follow the reference: ``Atom(atno, xyz)`` with xyz
in bohr (:258-297), ``Basis(atoms, name)`` over the named basis sets
(:300-343, build_basis src/Basis.jl:122-139), ``MolecularIntegrals()`` with
``all_twoe_ints`` and ``all_twoe_ints_4d`` (:442-482).  Unknown methods raise
ValueError as there; 'rys' runs Rys quadrature on the GPU (all_twoe_ints_rys,
src/Rys.jl:16-25; C ABI method MI_METHOD_RYS), the same integrals by the
other algorithm.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import basis as _basis
from .eri import default_context, _flat
from . import _lib

__all__ = ["Atom", "Basis", "MolecularIntegrals"]


class Atom:
    """Atom(atno, xyz): atomic number and (x, y, z) in bohr
    (python/molecular_integrals.py:258-297)."""

    def __init__(self, atno: int, xyz: Tuple[float, float, float]):
        self._atom = _basis.Atom(int(atno), np.asarray(xyz, dtype=float).reshape(3))

    @property
    def atno(self) -> int:
        return self._atom.atno

    @property
    def xyz(self) -> Tuple[float, float, float]:
        return tuple(float(v) for v in self._atom.xyz)

    def __repr__(self):
        return f"Atom(atno={self.atno}, xyz={self.xyz})"


class Basis:
    """Basis(atoms, name): build_basis over the named set ('sto3g', '6-31G',
    '6-31G**', 'cc-pvdz', matched without regard to case;
    python/molecular_integrals.py:300-343).  Indexing and iteration give the
    contracted functions (mi_amd.CGBF)."""

    def __init__(self, atoms: List[Atom], name: str = "sto3g"):
        from .basis_library import basis_data
        at = [a._atom if isinstance(a, Atom) else a for a in atoms]
        key = {k.lower(): k for k in basis_data}.get(name.lower())
        if key is None:
            raise KeyError(f"unknown basis set {name!r}; known: {sorted(basis_data)}")
        self.name = key
        self._basis = _basis.build_basis(at, key)

    @classmethod
    def from_basis(cls, b) -> "Basis":
        """Wrap an existing mi_amd.Basis (e.g. from basis_from_shells)."""
        obj = cls.__new__(cls)
        obj.name = None
        obj._basis = b
        return obj

    def __len__(self) -> int:
        return len(self._basis)

    def __getitem__(self, i: int):
        return self._basis[i]

    def __iter__(self):
        return iter(self._basis)

    def __repr__(self):
        return f"Basis(nbf={len(self)})"


def _expand_4d(ints_1d: np.ndarray, n: int) -> np.ndarray:
    """The (n, n, n, n) array of an iindex-ordered unique array: element
    [i, j, k, l] = ints_1d[iindex(i, j, k, l)] with the 0-based index algebra
    of python/molecular_integrals.py:467-482 (vectorised)."""
    i = np.arange(n, dtype=np.int64)
    hi, lo = np.maximum.outer(i, i), np.minimum.outer(i, i)
    pair = hi * (hi + 1) // 2 + lo               # [i, j] -> ij
    a, b = pair[:, :, None, None], pair[None, None, :, :]
    big, small = np.maximum(a, b), np.minimum(a, b)
    return ints_1d[big * (big + 1) // 2 + small]


class MolecularIntegrals:
    """High-level interface (python/molecular_integrals.py:346-482) over the
    GPU engine.  ``device`` selects the GPU (the reference has no such
    argument; its default, 0, matches a one-GPU box)."""

    _METHODS = ("default", "hgp", "rys")

    def __init__(self, device: int = 0):
        self._device = device
        self._own = None

    def _ctx(self):
        if self._device == 0:
            return default_context()
        if self._own is None:
            from .eri import Context
            self._own = Context(self._device)
        return self._own

    def all_twoe_ints(self, basis, method: str = "default") -> np.ndarray:
        """All unique (ij|kl) of the basis, 8-fold symmetry, iindex order
        (python/molecular_integrals.py:409-440; all_twoe_ints, src/ERI.jl:104-112)."""
        if method not in self._METHODS:
            raise ValueError(f"Unknown method: {method}")
        b = basis._basis if isinstance(basis, Basis) else basis
        m = _lib.MI_METHOD_RYS if method == "rys" else _lib.MI_METHOD_HGP
        return self._ctx().eri_unique(_flat(b), order=_lib.MI_ORDER_IINDEX, method=m)

    def all_twoe_ints_4d(self, basis, method: str = "default") -> np.ndarray:
        """All (ij|kl) as an (n, n, n, n) array (python/molecular_integrals.py:442-482)."""
        n = len(basis)
        return _expand_4d(self.all_twoe_ints(basis, method), n)
