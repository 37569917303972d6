"""Seeded synthetic workloads for the HGP ERI path.

All randomness comes from one ``numpy.random.Generator(PCG64(seed))`` per
workload, consumed in a fixed order (geometry first, then per atom, per shell:
exponents, then coefficients).  Coordinates are generated in Angstrom and
converted to bohr with the reference's factor (src/Atoms.jl:15).

Workloads
---------
ethane_minimal (16 functions)
    8 atoms in an ethane-like layout: heavy (Z=6) at (0,0,+-0.765) A; six light
    (Z=1) at radius 1.02 A, z=+1.16 A (phi 0/120/240 deg) and z=-1.16 A (phi
    60/180/300 deg); N(0, 0.02 A) jitter on every coordinate.  Heavy atoms
    carry s3 s3 p3, light atoms s3.  Exponents log-uniform in [0.1, 100],
    sorted descending; coefficients U(0.1, 1), scaled to unit 2-norm.
ethane_split_valence (60 functions)
    Same geometry recipe; heavy s6 s3 s1 p3 p1 d1, light s3 s1 p1; exponents
    log-uniform in [0.05, 5000] per shell, tight primitives first.
batch_molecules (512 molecules)
    Atom counts uniform in 3..24, each atom heavy with probability 0.3;
    positions uniform in a cube of edge 2.2*N^(1/3) A, each new atom redrawn
    until it is >= 0.9 A from every placed atom; split-valence recipe.
cluster_40 (920 functions)
    16 heavy + 24 light atoms, same placement rule; heavy s8 s3 s1 s1 p3 p1 p1
    d1 d1 f1, light s3 s1 s1 p1 p1 d1; exponents in [0.05, 5000]; density
    matrix symmetric with N(0,1)/n entries.
serving requests
    Poisson arrivals; atom count log-uniform in [2, 60]; split-valence or
    cluster recipe with probability 1/2 each; ERI request when nbf <= 120,
    otherwise a Coulomb/exchange request with a random symmetric density.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from .basis import BOHR_PER_ANGSTROM_DIVISOR, FlatBasis, basis_from_shells, flat_from_shells, ncart

# (L, nprim) per shell, per atom kind.
MINIMAL_HEAVY = [(0, 3), (0, 3), (1, 3)]
MINIMAL_LIGHT = [(0, 3)]
SPLIT_HEAVY = [(0, 6), (0, 3), (0, 1), (1, 3), (1, 1), (2, 1)]
SPLIT_LIGHT = [(0, 3), (0, 1), (1, 1)]
CLUSTER_HEAVY = [(0, 8), (0, 3), (0, 1), (0, 1), (1, 3), (1, 1), (1, 1), (2, 1), (2, 1), (3, 1)]
CLUSTER_LIGHT = [(0, 3), (0, 1), (0, 1), (1, 1), (1, 1), (2, 1)]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def ethane_geometry(rng: np.random.Generator) -> Tuple[np.ndarray, List[bool]]:
    pts = [(0.0, 0.0, 0.765), (0.0, 0.0, -0.765)]
    for z, phis in ((1.16, (0, 120, 240)), (-1.16, (60, 180, 300))):
        for phi in phis:
            a = math.radians(phi)
            pts.append((1.02 * math.cos(a), 1.02 * math.sin(a), z))
    xyz = np.array(pts, dtype=np.float64)
    xyz = xyz + rng.normal(0.0, 0.02, size=xyz.shape)
    return xyz, [True, True] + [False] * 6


def random_geometry(rng: np.random.Generator, nheavy_flags: Sequence[bool], min_dist: float = 0.9,
                    edge_per_cuberoot: float = 2.2) -> np.ndarray:
    n = len(nheavy_flags)
    edge = edge_per_cuberoot * n ** (1.0 / 3.0)
    pts = np.zeros((n, 3))
    for a in range(n):
        while True:
            p = rng.uniform(0.0, edge, size=3)
            if a == 0 or np.min(np.linalg.norm(pts[:a] - p, axis=1)) >= min_dist:
                pts[a] = p
                break
    return pts


def _shells_for(rng, xyz_ang, heavy_flags, heavy_recipe, light_recipe, lo, hi):
    spec = []
    for xyz, heavy in zip(xyz_ang, heavy_flags):
        for L, k in (heavy_recipe if heavy else light_recipe):
            ex = np.sort(np.exp(rng.uniform(math.log(lo), math.log(hi), size=k)))[::-1]
            co = rng.uniform(0.1, 1.0, size=k)
            co = co / np.linalg.norm(co)
            spec.append((np.asarray(xyz, dtype=np.float64) / BOHR_PER_ANGSTROM_DIVISOR, L, ex.tolist(), co.tolist()))
    return spec


def nbf_of(spec) -> int:
    return sum(ncart(L) for _, L, _, _ in spec)


@dataclass
class Workload:
    name: str
    spec: list          # [(xyz_bohr, L, expns, coefs)]
    heavy: List[bool]

    @property
    def nbf(self) -> int:
        return nbf_of(self.spec)

    def flat(self) -> FlatBasis:
        return flat_from_shells(self.spec)

    def basis(self):
        return basis_from_shells(self.spec)


def ethane_minimal(seed: int = 0) -> Workload:
    rng = _rng(seed)
    xyz, heavy = ethane_geometry(rng)
    return Workload("ethane_minimal_16bf", _shells_for(rng, xyz, heavy, MINIMAL_HEAVY, MINIMAL_LIGHT, 0.1, 100.0), heavy)


def ethane_split_valence(seed: int = 0) -> Workload:
    rng = _rng(seed)
    xyz, heavy = ethane_geometry(rng)
    return Workload("ethane_split_valence_60bf",
                    _shells_for(rng, xyz, heavy, SPLIT_HEAVY, SPLIT_LIGHT, 0.05, 5000.0), heavy)


def random_molecule(rng, natoms: int, heavy_recipe=SPLIT_HEAVY, light_recipe=SPLIT_LIGHT, p_heavy=0.3,
                    heavy_flags=None) -> Workload:
    if heavy_flags is None:
        heavy_flags = [bool(rng.random() < p_heavy) for _ in range(natoms)]
    xyz = random_geometry(rng, heavy_flags)
    spec = _shells_for(rng, xyz, heavy_flags, heavy_recipe, light_recipe, 0.05, 5000.0)
    return Workload(f"random_{natoms}atoms", spec, list(heavy_flags))


def batch_molecules(seed: int = 0, nmol: int = 512, natoms_lo: int = 3, natoms_hi: int = 24) -> List[Workload]:
    rng = _rng(seed)
    out = []
    for _ in range(nmol):
        n = int(rng.integers(natoms_lo, natoms_hi + 1))
        out.append(random_molecule(rng, n))
    return out


def cluster_40(seed: int = 0) -> Tuple[Workload, np.ndarray]:
    rng = _rng(seed)
    flags = [True] * 16 + [False] * 24
    wl = random_molecule(rng, 40, CLUSTER_HEAVY, CLUSTER_LIGHT, heavy_flags=flags)
    wl.name = "cluster_40atoms_920bf"
    n = wl.nbf
    d = rng.normal(0.0, 1.0, size=(n, n)) / n
    dens = 0.5 * (d + d.T)
    return wl, dens


@dataclass
class Request:
    rid: int
    arrival_s: float
    workload: Workload
    kind: str                  # "eri" or "jk"
    density: np.ndarray = None


def serving_requests(seed: int = 0, nreq: int = 64, rate_per_s: float = 200.0, natoms_lo: int = 2,
                     natoms_hi: int = 60, eri_max_nbf: int = 120) -> List[Request]:
    rng = _rng(seed)
    t = 0.0
    reqs = []
    for r in range(nreq):
        t += float(rng.exponential(1.0 / rate_per_s))
        n = int(round(math.exp(rng.uniform(math.log(natoms_lo), math.log(natoms_hi)))))
        if rng.random() < 0.5:
            wl = random_molecule(rng, n, SPLIT_HEAVY, SPLIT_LIGHT)
        else:
            wl = random_molecule(rng, n, CLUSTER_HEAVY, CLUSTER_LIGHT)
        if wl.nbf <= eri_max_nbf:
            reqs.append(Request(r, t, wl, "eri"))
        else:
            nb = wl.nbf
            d = rng.normal(0.0, 1.0, size=(nb, nb)) / nb
            reqs.append(Request(r, t, wl, "jk", 0.5 * (d + d.T)))
    return reqs
