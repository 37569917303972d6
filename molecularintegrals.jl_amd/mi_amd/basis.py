"""Host-side mirror of MolecularIntegrals.jl's basis model.

Mirrors the reference types and constructors (rpmuller/MolecularIntegrals.jl):

* ``PGBF`` / ``pgbf``            src/Basis.jl:14-33 (normalised on construction)
* ``CGBF`` / ``cgbf`` / ``addbf`` src/Basis.jl:53-83 (``addbf!`` renormalises)
* ``Shell``                      src/Basis.jl:147-153
* ``Basis`` / ``build_basis``    src/Basis.jl:122-139, :172-184
* ``atom``                       src/Atoms.jl:11-17 (Angstrom -> bohr by /0.52918)
* ``overlap``                    src/OneInts.jl:11-94 (used only for normalisation)
* ``shell_indices``, ``nao``, ``m2ao`` src/Basis.jl:269-321

``flatten(basis)`` turns a ``Basis`` into the plain arrays the C ABI takes
(``include/mi_hgp.h``): per shell the centre (bohr), L and primitive range; per
primitive the raw ``Shell.coefs`` and the PGBF norm of the shell's first
Cartesian component; per function ``bf_scale`` = CGBF norm times the ratio of
its component's PGBF norm to the first component's.  The integral the library
returns is then exactly ``coulomb(bfs[i],bfs[j],bfs[k],bfs[l])``
(src/ERI.jl:102), the package's tested value definition.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

BOHR_PER_ANGSTROM_DIVISOR = 0.52918  # src/Atoms.jl:15

# Cartesian component order per shell, src/Basis.jl:269-279 (note f: xyy, xzz, xyz).
shell_indices: Dict[int, List[Tuple[int, int, int]]] = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2), (1, 1, 1),
        (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3)],
    4: [(4, 0, 0), (3, 1, 0), (3, 0, 1), (2, 2, 0), (2, 1, 1), (2, 0, 2),
        (1, 3, 0), (1, 2, 1), (1, 1, 2), (1, 0, 3),
        (0, 4, 0), (0, 3, 1), (0, 2, 2), (0, 1, 3), (0, 0, 4)],
}
llabel = {0: "s", 1: "p", 2: "d", 3: "f", 4: "g", 5: "h"}
lvalue = {**{v: k for k, v in llabel.items()}, **{v.upper(): k for k, v in llabel.items()}}


def nao(l: int) -> int:
    """Cumulative number of Cartesian AOs for shells 0..l (src/Basis.jl:320-321)."""
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def m2ao(ijk: Sequence[int]) -> int:
    """1-based cumulative AO index of a power triple (src/Basis.jl:295-305)."""
    I, J, K = (int(x) for x in ijk)
    L = I + J + K
    return nao(L - 1) + shell_indices[L].index((I, J, K)) + 1


def factorial2(n: int) -> int:
    """prod(n:-2:1); factorial2(-1) == 1 (src/Utils.jl:5)."""
    p = 1
    for k in range(n, 0, -2):
        p *= k
    return p


# --------------------------------------------------------------------------
# Overlap, src/OneInts.jl:11-94 (only what normalisation needs).
# --------------------------------------------------------------------------
def _binomial_prefactor(s: int, ia: int, ib: int, xpa: float, xpb: float) -> float:
    total = 0.0
    for t in range(s + 1):
        if s - ia <= t <= ib:
            total += math.comb(ia, s - t) * xpa ** (ia - s + t) * math.comb(ib, t) * xpb ** (ib - t)
    return total


def _overlap1d(la: int, lb: int, ax: float, bx: float, gamma: float) -> float:
    total = 0.0
    for i in range((la + lb) // 2 + 1):
        total += _binomial_prefactor(2 * i, la, lb, ax, bx) * factorial2(2 * i - 1) / (2 * gamma) ** i
    return total


def overlap_prim(aexpn, axyz, aI, aJ, aK, bexpn, bxyz, bI, bJ, bK) -> float:
    """Unnormalised primitive overlap (src/OneInts.jl:41-52)."""
    gamma = aexpn + bexpn
    axyz = np.asarray(axyz, dtype=float)
    bxyz = np.asarray(bxyz, dtype=float)
    p = (aexpn / gamma) * axyz + (bexpn / gamma) * bxyz
    pa, pb = p - axyz, p - bxyz
    rab2 = float(np.dot(axyz - bxyz, axyz - bxyz))
    pre = (math.pi / gamma) ** 1.5 * math.exp(-aexpn * bexpn * rab2 / gamma)
    return (pre * _overlap1d(aI, bI, pa[0], pb[0], gamma) * _overlap1d(aJ, bJ, pa[1], pb[1], gamma)
            * _overlap1d(aK, bK, pa[2], pb[2], gamma))


@dataclass
class PGBF:
    """Primitive Cartesian Gaussian ``norm*(x-x0)^I (y-y0)^J (z-z0)^K exp(-expn r^2)``."""
    expn: float
    xyz: np.ndarray
    I: int
    J: int
    K: int
    norm: float = 1.0

    def __call__(self, x, y, z):  # amplitude, src/Basis.jl:35-40
        dx, dy, dz = self.xyz - np.array([x, y, z], dtype=float)
        return self.norm * dx ** self.I * dy ** self.J * dz ** self.K * math.exp(-self.expn * (dx * dx + dy * dy + dz * dz))


def overlap(a, b) -> float:
    """Normalised overlap of two PGBFs or two CGBFs (src/OneInts.jl:11-22)."""
    if isinstance(a, PGBF):
        return a.norm * b.norm * overlap_prim(a.expn, a.xyz, a.I, a.J, a.K, b.expn, b.xyz, b.I, b.J, b.K)
    s = 0.0
    for ca, pa in zip(a.coefs, a.pgbfs):
        for cb, pb in zip(b.coefs, b.pgbfs):
            s += ca * cb * overlap(pa, pb)
    return a.norm * b.norm * s


def pgbf(expn, x=0.0, y=0.0, z=0.0, I=0, J=0, K=0, norm=1.0) -> PGBF:
    """Normalised PGBF (src/Basis.jl:29-33, normalize! :42-44)."""
    p = PGBF(float(expn), np.array([x, y, z], dtype=float), int(I), int(J), int(K), float(norm))
    p.norm /= math.sqrt(overlap(p, p))
    return p


@dataclass
class CGBF:
    xyz: np.ndarray
    I: int
    J: int
    K: int
    norm: float = 1.0
    pgbfs: List[PGBF] = field(default_factory=list)
    coefs: List[float] = field(default_factory=list)

    def __call__(self, x, y, z):
        return self.norm * sum(c * p(x, y, z) for c, p in zip(self.coefs, self.pgbfs))


def cgbf(x=0.0, y=0.0, z=0.0, I=0, J=0, K=0) -> CGBF:
    return CGBF(np.array([x, y, z], dtype=float), int(I), int(J), int(K), 1.0, [], [])


def addbf(cbf: CGBF, expn, coef) -> None:
    """``addbf!``: append a normalised primitive and renormalise (src/Basis.jl:79-83)."""
    cbf.pgbfs.append(pgbf(expn, *cbf.xyz, cbf.I, cbf.J, cbf.K))
    cbf.coefs.append(float(coef))
    cbf.norm /= math.sqrt(overlap(cbf, cbf))


@dataclass
class Shell:
    xyz: np.ndarray
    L: int
    expns: List[float]
    coefs: List[float]
    cgbfs: List[CGBF] = field(default_factory=list)


@dataclass
class Basis:
    cgbfs: List[CGBF]
    shells: List[Shell]
    ishell: List[int]  # 1-based shell of each function
    mshell: List[int]  # 1-based position within the shell

    def __len__(self):
        return len(self.cgbfs)

    def __getitem__(self, i):
        return self.cgbfs[i]

    def __iter__(self):
        return iter(self.cgbfs)


@dataclass
class Atom:
    atno: int
    xyz: np.ndarray  # bohr


def atom(atno: int, x: float, y: float, z: float, units: str = "Angstrom") -> Atom:
    xyz = np.array([x, y, z], dtype=float)
    if units == "Angstrom":
        xyz = xyz / BOHR_PER_ANGSTROM_DIVISOR
    return Atom(int(atno), xyz)


def basis_from_shells(shells_spec: Sequence[Tuple[Sequence[float], int, Sequence[float], Sequence[float]]]) -> Basis:
    """Build a ``Basis`` from ``(xyz_bohr, L, expns, coefs)`` tuples, exactly as
    ``build_basis`` does from library data (src/Basis.jl:100-139)."""
    shs: List[Shell] = []
    for xyz, L, expns, coefs in shells_spec:
        xyz = np.asarray(xyz, dtype=float)
        bfs = []
        for (I, J, K) in shell_indices[L]:
            cbf = cgbf(*xyz, I, J, K)
            for e, c in zip(expns, coefs):
                addbf(cbf, e, c)
            bfs.append(cbf)
        shs.append(Shell(xyz, int(L), [float(e) for e in expns], [float(c) for c in coefs], bfs))
    cg, ish, msh = [], [], []
    for si, sh in enumerate(shs, start=1):
        for mi, (I, J, K) in enumerate(shell_indices[sh.L], start=1):
            cbf = cgbf(*sh.xyz, I, J, K)
            for e, c in zip(sh.expns, sh.coefs):
                addbf(cbf, e, c)
            cg.append(cbf)
            ish.append(si)
            msh.append(mi)
    return Basis(cg, shs, ish, msh)


def build_basis(atoms: Sequence[Atom], name: str = "sto3g") -> Basis:
    """``build_basis(atoms, name)`` over the named-basis library (src/Basis.jl:122-139)."""
    from .basis_library import basis_data

    data = basis_data[name]
    spec = []
    for at in atoms:
        for sym, primlist in data[at.atno]:
            spec.append((at.xyz, lvalue[sym], [e for e, _ in primlist], [c for _, c in primlist]))
    return basis_from_shells(spec)


# --------------------------------------------------------------------------
# Flattening for the C ABI.
# --------------------------------------------------------------------------
@dataclass
class FlatBasis:
    shell_xyz: np.ndarray      # (nshell, 3) float64, bohr
    shell_l: np.ndarray        # (nshell,) int32
    shell_prim_off: np.ndarray # (nshell+1,) int32
    prim_exp: np.ndarray       # (nprim,) float64
    prim_coef: np.ndarray      # (nprim,) float64, raw Shell.coefs
    prim_norm: np.ndarray      # (nprim,) float64, PGBF norm of the shell's first component
    bf_scale: np.ndarray       # (nbf,) float64

    @property
    def nshell(self) -> int:
        return int(self.shell_l.shape[0])

    @property
    def nbf(self) -> int:
        return int(self.bf_scale.shape[0])

    def arrays(self):
        return (np.ascontiguousarray(self.shell_xyz, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(self.shell_l, dtype=np.int32),
                np.ascontiguousarray(self.shell_prim_off, dtype=np.int32),
                np.ascontiguousarray(self.prim_exp, dtype=np.float64),
                np.ascontiguousarray(self.prim_coef, dtype=np.float64),
                np.ascontiguousarray(self.prim_norm, dtype=np.float64),
                np.ascontiguousarray(self.bf_scale, dtype=np.float64))


def flatten(bfs: Basis) -> FlatBasis:
    """Flatten a ``Basis`` into the C-ABI arrays, reading the norms the objects hold."""
    xyz, ls, off, ex, co, pn, sc = [], [], [0], [], [], [], []
    for sh in bfs.shells:
        xyz.append(np.asarray(sh.xyz, dtype=float))
        ls.append(sh.L)
        first = sh.cgbfs[0] if sh.cgbfs else None
        for p, (e, c) in enumerate(zip(sh.expns, sh.coefs)):
            ex.append(e)
            co.append(c)
            pn.append(first.pgbfs[p].norm if first is not None else pgbf(e, 0, 0, 0, *shell_indices[sh.L][0]).norm)
        off.append(len(ex))
    first_norm0 = {}
    for si, sh in enumerate(bfs.shells, start=1):
        first_norm0[si] = (sh.cgbfs[0].pgbfs[0].norm if sh.cgbfs
                           else pgbf(sh.expns[0], 0, 0, 0, *shell_indices[sh.L][0]).norm)
    for f, cbf in enumerate(bfs.cgbfs):
        si = bfs.ishell[f]
        sc.append(cbf.norm * cbf.pgbfs[0].norm / first_norm0[si])
    return FlatBasis(np.array(xyz, dtype=np.float64).reshape(-1, 3), np.array(ls, dtype=np.int32),
                     np.array(off, dtype=np.int32), np.array(ex, dtype=np.float64),
                     np.array(co, dtype=np.float64), np.array(pn, dtype=np.float64),
                     np.array(sc, dtype=np.float64))


def flat_from_shells(shells_spec) -> FlatBasis:
    """Vectorised equivalent of ``flatten(basis_from_shells(spec))``.

    Uses the closed forms the reference's overlap-based normalisation reduces to
    for a single-centre function: PGBF norm n = (2a/pi)^{3/4} (4a)^{L/2} /
    sqrt(prod (2I-1)!!), CGBF norm 1/sqrt(sum_pq c_p c_q n_p n_q S_pq) with the
    same-centre overlap S_pq = (pi/(a_p+a_q))^{3/2} prod (2I-1)!!/(2(a_p+a_q))^I.
    Used for large synthetic workloads; tests pin it against ``flatten``.
    """
    xyz, ls, off, ex, co, pn, sc = [], [], [0], [], [], [], []
    for sxyz, L, expns, coefs in shells_spec:
        e = np.asarray(expns, dtype=np.float64)
        c = np.asarray(coefs, dtype=np.float64)
        xyz.append(np.asarray(sxyz, dtype=np.float64))
        ls.append(int(L))
        comps = shell_indices[L]
        # radial part of the PGBF norm, first component's angular factor
        f0 = 1.0 / math.sqrt(np.prod([factorial2(2 * q - 1) for q in comps[0]]))
        nrad = (2 * e / math.pi) ** 0.75 * (4 * e) ** (L / 2.0)
        n0 = nrad * f0
        ex.extend(e.tolist())
        co.extend(c.tolist())
        pn.extend(n0.tolist())
        off.append(len(ex))
        g = e[:, None] + e[None, :]
        for (I, J, K) in comps:
            dfac = float(factorial2(2 * I - 1) * factorial2(2 * J - 1) * factorial2(2 * K - 1))
            fcomp = 1.0 / math.sqrt(dfac)
            n = nrad * fcomp
            S = (math.pi / g) ** 1.5 * dfac / (2 * g) ** L
            s = float(np.sum(c[:, None] * c[None, :] * n[:, None] * n[None, :] * S))
            sc.append((1.0 / math.sqrt(s)) * fcomp / f0)
    return FlatBasis(np.array(xyz, dtype=np.float64).reshape(-1, 3), np.array(ls, dtype=np.int32),
                     np.array(off, dtype=np.int32), np.array(ex, dtype=np.float64),
                     np.array(co, dtype=np.float64), np.array(pn, dtype=np.float64),
                     np.array(sc, dtype=np.float64))


def sub_basis(fb: FlatBasis, shells) -> FlatBasis:
    """The basis of the given shells of ``fb`` (in that order), with ``fb``'s
    primitive norms and function scales unchanged, so every integral over its
    functions equals the same integral of ``fb``.  Used to sample integrals of
    bases whose whole unique array is too large to hold (a 920-function
    cluster has 9e10 unique integrals)."""
    shells = [int(s) for s in shells]
    nbf_shell = [ncart(int(L)) for L in fb.shell_l]
    bf0 = np.concatenate([[0], np.cumsum(nbf_shell)]).astype(np.int64)
    off, ex, co, pn, sc = [0], [], [], [], []
    for s in shells:
        p0, p1 = int(fb.shell_prim_off[s]), int(fb.shell_prim_off[s + 1])
        ex.append(fb.prim_exp[p0:p1])
        co.append(fb.prim_coef[p0:p1])
        pn.append(fb.prim_norm[p0:p1])
        off.append(off[-1] + p1 - p0)
        sc.append(fb.bf_scale[bf0[s]:bf0[s + 1]])
    return FlatBasis(shell_xyz=np.ascontiguousarray(fb.shell_xyz[shells], dtype=np.float64),
                     shell_l=np.ascontiguousarray(fb.shell_l[shells], dtype=np.int32),
                     shell_prim_off=np.asarray(off, dtype=np.int32),
                     prim_exp=np.concatenate(ex), prim_coef=np.concatenate(co), prim_norm=np.concatenate(pn),
                     bf_scale=np.concatenate(sc))


def shell_functions(fb: FlatBasis, shells) -> np.ndarray:
    """Indices (0-based) in ``fb`` of the functions of ``sub_basis(fb, shells)``, in order."""
    nbf_shell = [ncart(int(L)) for L in fb.shell_l]
    bf0 = np.concatenate([[0], np.cumsum(nbf_shell)]).astype(np.int64)
    return np.concatenate([np.arange(bf0[int(s)], bf0[int(s) + 1]) for s in shells])
