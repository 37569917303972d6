"""Algorithmic FLOP model of the HGP path, from the reference's loop trip counts.

Per canonical shell quartet of class (la,lb|lc,ld), with K primitive quartets,
amax = la+lb, cmax = lc+ld, mmax = amax+cmax+1:

    F = K * [S + 4*mmax + V(amax,cmax) + 2*nao(amax)*nao(cmax)] + H

* S = 60: quartet geometry, prefactor, W and T (exp/sqrt/div count 1 each).
* 4*mmax: the Boys function.
* V: trip counts of ``vrr!`` (rpmuller/MolecularIntegrals.jl src/HGP.jl:169-240)
  weighted per statement: a-build 3 ops, 7 when a-1_i exists; c-build 3,
  +2 when a-1_i exists, +4 when c-1_i exists; plus mmax for the [0|0]^(m)
  scaling.
* 2*nao(amax)*nao(cmax): the contraction in ``cvrr`` (src/HGP.jl:62-64).
* H = 2 x the HRR entries the target block needs (one FMA each).

The same numbers drive ``roofline.achieved`` in bench.py.
"""
from __future__ import annotations

import functools

from .hgp_tables import AO2M, HrrGraph, lval, nao, sdir, sidx

S_SETUP = 60


@functools.lru_cache(maxsize=None)
def vrr_ops(amax: int, cmax: int) -> int:
    mmax = amax + cmax + 1
    ops = mmax
    for aplus in range(1, nao(amax)):
        i = sdir(aplus)
        a = sidx(aplus, i)
        lim = mmax - lval(aplus)
        aminus = sidx(a, i)
        ops += max(lim, 0) * (7 if aminus >= 0 else 3)
    for cplus in range(1, nao(cmax)):
        i = sdir(cplus)
        c = sidx(cplus, i)
        cminus = sidx(c, i)
        for a in range(nao(amax)):
            lim = mmax - lval(cplus) - lval(a)
            if lim <= 0:
                continue
            aminus = sidx(a, i)
            ops += lim * (3 + (2 if aminus >= 0 else 0) + (4 if cminus >= 0 else 0))
    return ops


@functools.lru_cache(maxsize=None)
def hrr_ops(la: int, lb: int, lc: int, ld: int) -> int:
    h = HrrGraph(la, lb, lc, ld)
    return 2 * sum(1 for n in h.nodes if h.deps(n)[0] != "acc")


def per_prim_quartet(la, lb, lc, ld) -> int:
    amax, cmax = la + lb, lc + ld
    mmax = amax + cmax + 1
    return S_SETUP + 4 * mmax + vrr_ops(amax, cmax) + 2 * nao(amax) * nao(cmax)


def class_flops(cls, n_shell_quartets: int, n_prim_quartets: int) -> int:
    la, lb, lc, ld = cls
    return n_prim_quartets * per_prim_quartet(la, lb, lc, ld) + n_shell_quartets * hrr_ops(la, lb, lc, ld)
