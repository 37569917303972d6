"""Cartesian AO tables and the pruned HGP recurrence graphs.

Shared by the code generator (csrc/gen_hgp.py) and the FLOP model
(mi_amd/flops.py).  Restates the reference's tables (rpmuller/
MolecularIntegrals.jl src/Basis.jl:264-385: shell_indices, m2ao, ao2m, nao,
shift_index, shift_direction, shell_number, shift_index_plus) and the
recurrence choices of vrr! (src/HGP.jl:135-242) and chrr (src/HGP.jl:83-116).
"""
from __future__ import annotations

import sys
from collections import OrderedDict

LMAX_SHELL = 3           # shells up to f
LMAX_TAB = 2 * 3 + 1     # AO tables cover HRR intermediates of f shells too

# ---------------------------------------------------------------- AO tables


def shell_components(L):
    if L == 3:
        return [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2), (1, 1, 1),
                (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3)]
    return [(I, J, L - I - J) for I in range(L, -1, -1) for J in range(L - I, -1, -1)]


def nao(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def ncart(l):
    return (l + 1) * (l + 2) // 2


AO2M = [c for L in range(LMAX_TAB + 1) for c in shell_components(L)]
M2AO = {c: i for i, c in enumerate(AO2M)}


def lval(a):
    return sum(AO2M[a])


def sdir(a):
    m = AO2M[a]
    best = 0
    for i in (1, 2):
        if m[i] > m[best]:
            best = i
    return best


def sidx(a, i):
    m = list(AO2M[a])
    if m[i] == 0:
        return -1
    m[i] -= 1
    return M2AO[tuple(m)]


def sidx_plus(a, i):
    m = list(AO2M[a])
    m[i] += 1
    return M2AO[tuple(m)]


def shell_aos(L):
    return list(range(nao(L - 1), nao(L)))


# ---------------------------------------------------------------- classes


def pair_class_index(la, lb, lmax=LMAX_SHELL):
    """Index of the pair class (L_A, L_B) in pair_classes(lmax)."""
    if la >= lb:
        return la * (la + 1) // 2 + lb
    return (lmax + 1) * (lmax + 2) // 2 + lb * (lb - 1) // 2 + la


def pair_classes(lmax=LMAX_SHELL):
    """Pair classes (L_A, L_B): the canonical L_A >= L_B ones first (index
    L_A (L_A+1)/2 + L_B, the reference's chrr shell order up to a swap), then
    the swapped L_A < L_B ones.  A shell pair is swapped when its HRR would
    otherwise cancel many digits: a diffuse high-L shell A far from a tight
    low-L shell B puts P next to B, so the (e0| terms grow like |AB|^e while
    the result does not (see mi_host.cpp, pair orientation)."""
    normal = [(la, lb) for la in range(lmax + 1) for lb in range(la + 1)]
    swapped = [(la, lb) for lb in range(1, lmax + 1) for la in range(lb)]
    return normal + swapped


def is_swapped_class(c):
    """A quartet class with a swapped pair (L_A < L_B or L_C < L_D)."""
    return c[0] < c[1] or c[2] < c[3]


def quartet_classes(lmax=LMAX_SHELL):
    pcs = pair_classes(lmax)
    out = []
    for X, (la, lb) in enumerate(pcs):
        for Y, (lc, ld) in enumerate(pcs):
            if Y <= X:
                out.append((la, lb, lc, ld))
    return out


def cname(c):
    return "".join("spdfghik"[x] for x in c)


# ---------------------------------------------------------------- VRR graph


class VrrGraph:
    """Pruned VRR dependency graph for a class; nodes (e, f, m)."""

    def __init__(self, la, lb, lc, ld):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.amax, self.cmax = la + lb, lc + ld
        self.mmax = self.amax + self.cmax + 1
        self.targets = [(e, f) for e in range(nao(la - 1), nao(self.amax))
                        for f in range(nao(lc - 1), nao(self.cmax))]
        # max m needed per (e,f), propagated from targets (m=0)
        maxm = {}
        order = []  # (e,f) in reverse-topological discovery

        def need(e, f, m):
            if (e, f) in maxm and maxm[(e, f)] >= m:
                return
            maxm[(e, f)] = max(maxm.get((e, f), -1), m)
            for (x, mm) in self.deps(e, f, m):
                need(x[0], x[1], mm)

        sys.setrecursionlimit(100000)
        for (e, f) in self.targets:
            need(e, f, 0)
        self.maxm = maxm
        # topological order: by level then by (e,f)
        self.nodes = sorted(maxm.keys(), key=lambda ef: (self.level(*ef), ef[1], ef[0]))

    def level(self, e, f):
        return lval(e) if f == 0 else self.amax + lval(f)

    def deps(self, e, f, m):
        """Dependencies of [e|f]^(m) as ((e',f'), m') pairs."""
        out = []
        if f == 0:
            if e == 0:
                return out
            i = sdir(e)
            e1 = sidx(e, i)
            e2 = sidx(e1, i)
            out += [((e1, 0), m), ((e1, 0), m + 1)]
            if e2 >= 0:
                out += [((e2, 0), m), ((e2, 0), m + 1)]
            return out
        i = sdir(f)
        f1 = sidx(f, i)
        f2 = sidx(f1, i)
        e1 = sidx(e, i)
        out += [((e, f1), m), ((e, f1), m + 1)]
        if e1 >= 0:
            out += [((e1, f1), m + 1)]
        if f2 >= 0:
            out += [((e, f2), m), ((e, f2), m + 1)]
        return out

    def recurrence(self, e, f):
        """Description of how [e|f] is formed (shared by R and W emitters)."""
        if f == 0:
            if e == 0:
                return ("F",)
            i = sdir(e)
            e1 = sidx(e, i)
            e2 = sidx(e1, i)
            return ("A", i, e1, e2, AO2M[e1][i])  # n1 = N_i(e-1_i)
        i = sdir(f)
        f1 = sidx(f, i)
        f2 = sidx(f1, i)
        e1 = sidx(e, i)
        return ("C", i, f1, f2, e1, AO2M[e][i], AO2M[f1][i])  # N_i(e), N_i(f-1_i)

    def count_entries(self):
        return sum(m + 1 for m in self.maxm.values())


# ---------------------------------------------------------------- HRR graph


class HrrGraph:
    """Pruned HRR graph; nodes (a, b, c, d) with cumulative AO indices."""

    def __init__(self, la, lb, lc, ld):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.targets = [(a, b, c, d) for a in shell_aos(la) for b in shell_aos(lb)
                        for c in shell_aos(lc) for d in shell_aos(ld)]
        seen = OrderedDict()

        def visit(n):
            if n in seen:
                return
            for x in self.deps(n)[1:3]:
                if x is not None:
                    visit(x)
            seen[n] = True

        for t in self.targets:
            visit(t)
        self.nodes = list(seen.keys())  # topological (post-order)

    @staticmethod
    def deps(n):
        a, b, c, d = n
        if d != 0:
            j = sdir(d)
            d1 = sidx(d, j)
            return ("ket", (a, b, sidx_plus(c, j), d1), (a, b, c, d1), j)
        if b != 0:
            j = sdir(b)
            b1 = sidx(b, j)
            return ("bra", (sidx_plus(a, j), b1, c, 0), (a, b1, c, 0), j)
        return ("acc", None, None, None)

    def level(self, n):
        a, b, c, d = n
        if d != 0:
            return 1 + self.lb + lval(d)
        if b != 0:
            return lval(b)
        return 0




# Fusion groups of the register-kernel classes by total angular momentum (a
# proxy for their VGPR use: about 70-125 for L <= 2, 130-205 for L = 3,
# 200-256 for L = 4, the 256 cap above), the heaviest split into d and f
# classes.  A fused kernel takes the VGPRs of its largest class body, so light
# classes keep their occupancy.
NRGROUP = 5


def register_group(c):
    L = sum(c)
    if L <= 2:
        return 0
    if L == 3:
        return 1
    if L == 4:
        return 2
    return 3 if max(c) <= 2 else 4
