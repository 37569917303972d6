"""Writes tests/golden/cluster40_outliers.npz: exact values of the hardest
integrals of the 40-atom, 920-function cluster (f shells), for pinning the
binary128 oracle and checking the GPU at the parity bar.

Sample: the 1,820 random sets of 4 distinct shells of cluster_40(seed 0)
drawn with PCG64(77) -- the shell sets bench.py's eri_spot_check uses.  Each
set is a sub-basis with the cluster's own norms (mi_amd.sub_basis), so its
integrals are the cluster's.

Selection (12.35M integrals in all), against the binary128 oracle with the
Boys function exact for every T:
  kind 0  the 150 integrals where the FP64 restatement of the reference
          algorithm (oracle.eri_unique_hgp) is furthest from it: far-apart
          tight/diffuse shell pairs whose HRR cancels 5-9 digits;
  kind 1  the next 50 where the x87 80-bit restatement is furthest;
  kind 2  the next 50 where the reference's T >= 30 Boys branch (farray,
          src/Boys.jl:21-26, binary128 with literal D5) is furthest;
  kind 3  200 random integrals (at most 64 primitive quartets each).

Exact value: an independent restatement of the THO/Huzinaga primitive ERI
(rpmuller/MolecularIntegrals.jl src/ERI.jl:15-46, Barray/Bterm :62-94,
binomial_prefactor src/OneInts.jl:93) in mpmath at 40 digits, with the Boys
function exact for every T (Fgamma = Fm = gamma_inc form, src/Boys.jl:29-34;
here the lower incomplete gamma of mpmath), contracted with the flattened
basis' coefficients, primitive norms and function scales (src/ERI.jl:55-60,102).

Run in this repository's container only (it needs the oracle library and
mpmath; about 5 minutes on 8 cores):  python tests/golden/make_cluster_outliers.py
"""
import os
import sys
from multiprocessing import Pool

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "molecularintegrals.jl_amd"))

from oracle import oracle as O  # noqa: E402
from mi_amd import configs, sub_basis  # noqa: E402
from mi_amd.hgp_tables import shell_components  # noqa: E402

NSETS = 1820
DPS = 40


def cluster_sets():
    wl, _ = configs.cluster_40(0)
    fb = wl.flat()
    rng = np.random.Generator(np.random.PCG64(77))
    return fb, [rng.choice(fb.nshell, 4, replace=False) for _ in range(NSETS)]


def chrr_quartet(n, slot):
    """(i, j, k, l), 0-based functions, of 0-based chrr slot `slot` (the
    enumeration of iiterator, src/Utils.jl:17, src/Basis.jl:203)."""
    M = n * (n + 1) // 2
    # slot = (ij-1)(M+1) - ij(ij-1)/2 + kl - ij with ij <= kl, 1-based pair indices
    ij = 1
    while True:
        first = (ij - 1) * (M + 1) - ij * (ij - 1) // 2  # 0-based slot of (ij, kl = ij)
        if slot < first + (M - ij + 1):
            kl = ij + slot - first
            break
        ij += 1

    def unpair(p):
        i = 1
        while i * (i + 1) // 2 < p:
            i += 1
        return i - 1, p - i * (i - 1) // 2 - 1

    i, j = unpair(ij)
    k, l = unpair(kl)
    return i, j, k, l


# ---------------------------------------------------------------- mpmath THO

def _binomial_prefactor(s, ia, ib, xpa, xpb):
    return mpmath.fsum(mpmath.binomial(ia, s - t) * xpa ** (ia - s + t) * mpmath.binomial(ib, t) * xpb ** (ib - t)
                       for t in range(s + 1) if s - ia <= t <= ib)


def _fact_ratio2(a, b):
    return mpmath.factorial(a) / mpmath.factorial(b) / mpmath.factorial(a - 2 * b)


def _barray(l1, l2, l3, l4, p, a, b, q, c, d, g1, g2, delta):
    B = [mpmath.mpf(0)] * (l1 + l2 + l3 + l4 + 1)
    for i1 in range(l1 + l2 + 1):
        for i2 in range(l3 + l4 + 1):
            for r1 in range(i1 // 2 + 1):
                for r2 in range(i2 // 2 + 1):
                    for u in range((i1 + i2) // 2 - r1 - r2 + 1):
                        e = i1 + i2 - 2 * (r1 + r2)
                        t = (_binomial_prefactor(i1, l1, l2, p - a, p - b) * _fact_ratio2(i1, r1) * (4 * g1) ** (r1 - i1)
                             * (-1) ** i2 * _binomial_prefactor(i2, l3, l4, q - c, q - d) * _fact_ratio2(i2, r2)
                             * (4 * g2) ** (r2 - i2) * (-1) ** u * _fact_ratio2(e, u) * (q - p) ** (e - 2 * u)
                             / delta ** (e - u))
                        B[e - u] += t
    return B


def _boys(m, x):
    if x == 0:
        return mpmath.mpf(1) / (2 * m + 1)
    a = m + mpmath.mpf(1) / 2
    return mpmath.gammainc(a, 0, x) / (2 * x ** a)


def coulomb_prim(ea, A, Ia, eb, B, Ib, ec, C, Ic, ed, D, Id):
    """THO primitive (ab|cd) without norms, src/ERI.jl:15-46, in mpmath."""
    g1, g2 = ea + eb, ec + ed
    P = [(ea * A[k] + eb * B[k]) / g1 for k in range(3)]
    Q = [(ec * C[k] + ed * D[k]) / g2 for k in range(3)]
    rab2 = sum((A[k] - B[k]) ** 2 for k in range(3))
    rcd2 = sum((C[k] - D[k]) ** 2 for k in range(3))
    rpq2 = sum((P[k] - Q[k]) ** 2 for k in range(3))
    delta = (1 / g1 + 1 / g2) / 4
    Bs = [_barray(Ia[k], Ib[k], Ic[k], Id[k], P[k], A[k], B[k], Q[k], C[k], D[k], g1, g2, delta) for k in range(3)]
    x = rpq2 / (4 * delta)
    F = {}
    s = mpmath.mpf(0)
    for I, bx in enumerate(Bs[0]):
        for J, by in enumerate(Bs[1]):
            for K, bz in enumerate(Bs[2]):
                n = I + J + K
                if n not in F:
                    F[n] = _boys(n, x)
                s += bx * by * bz * F[n]
    return 2 * mpmath.pi ** mpmath.mpf(2.5) / (g1 * g2 * mpmath.sqrt(g1 + g2)) * mpmath.exp(-ea * eb * rab2 / g1) \
        * mpmath.exp(-ec * ed * rcd2 / g2) * s


def exact_integral(args):
    """Contracted, normalised (ij|kl) of a sub-basis (src/ERI.jl:55-60,102)."""
    sb, (i, j, k, l) = args
    mpmath.mp.dps = DPS
    nc = (sb.shell_l + 1) * (sb.shell_l + 2) // 2
    bf0 = np.concatenate([[0], np.cumsum(nc)])
    total = mpmath.mpf(0)
    parts = []
    for f in (i, j, k, l):
        s = int(np.searchsorted(bf0, f, side="right") - 1)
        comp = shell_components(int(sb.shell_l[s]))[f - bf0[s]]
        prims = [(mpmath.mpf(float(sb.prim_exp[p])), mpmath.mpf(float(sb.prim_coef[p])) * mpmath.mpf(float(sb.prim_norm[p])))
                 for p in range(sb.shell_prim_off[s], sb.shell_prim_off[s + 1])]
        parts.append(([mpmath.mpf(float(x)) for x in sb.shell_xyz[s]], comp, prims, mpmath.mpf(float(sb.bf_scale[f]))))
    (A, Ia, pa, sa), (B, Ib, pb, sb_), (C, Ic, pc, sc), (D, Id, pd, sd) = parts
    for ea, ca in pa:
        for eb, cb in pb:
            for ec, cc in pc:
                for ed, cd in pd:
                    total += ca * cb * cc * cd * coulomb_prim(ea, A, Ia, eb, B, Ib, ec, C, Ic, ed, D, Id)
    return float(total * sa * sb_ * sc * sd)


def nprim_quartet(sb, q):
    npr = np.diff(sb.shell_prim_off)
    nc = (sb.shell_l + 1) * (sb.shell_l + 2) // 2
    bf0 = np.concatenate([[0], np.cumsum(nc)])
    return int(np.prod([npr[int(np.searchsorted(bf0, f, side="right") - 1)] for f in q]))


def main():
    fb, sets = cluster_sets()
    subs = [sub_basis(fb, s) for s in sets]
    err64, err80, errb, sizes = [], [], [], []
    for sb in subs:
        q = O.eri_unique_hgp_quad(sb, order=O.ORDER_CHRR, accurate=True, nthreads=8)
        err64.append(np.abs(O.eri_unique_hgp(sb, order=O.ORDER_CHRR, nthreads=8) - q))
        err80.append(np.abs(O.eri_unique_hgp_ext(sb, order=O.ORDER_CHRR, nthreads=8) - q))
        errb.append(np.abs(O.eri_unique_hgp_quad(sb, order=O.ORDER_CHRR, accurate=False, nthreads=8) - q))
        sizes.append(q.size)
    off = np.concatenate([[0], np.cumsum(sizes)])
    chosen, kinds = [], []
    for kind, (errs, upto) in enumerate(((err64, 150), (err80, 200), (errb, 250))):
        for g in np.argsort(-np.concatenate(errs)):
            if len(kinds) == upto:
                break
            if int(g) not in chosen:
                chosen.append(int(g))
                kinds.append(kind)
    rng = np.random.Generator(np.random.PCG64(2024))
    while len(chosen) < 450:
        g = int(rng.integers(0, off[-1]))
        s = int(np.searchsorted(off, g, side="right") - 1)
        q = chrr_quartet(subs[s].nbf, g - off[s])
        if g in chosen or nprim_quartet(subs[s], q) > 64:
            continue
        chosen.append(g)
        kinds.append(3)
    set_id = np.array([int(np.searchsorted(off, g, side="right") - 1) for g in chosen], dtype=np.int32)
    slot = np.array([g - off[s] for g, s in zip(chosen, set_id)], dtype=np.int64)
    jobs = [(subs[s], chrr_quartet(subs[s].nbf, int(t))) for s, t in zip(set_id, slot)]
    with Pool(8) as p:
        exact = np.array(p.map(exact_integral, jobs, chunksize=1))
    used = np.unique(set_id)
    remap = {int(s): k for k, s in enumerate(used)}
    np.savez(os.path.join(HERE, "cluster40_outliers.npz"),
             shells=np.array([sets[int(s)] for s in used], dtype=np.int32),
             set_index=used.astype(np.int32),
             set_id=np.array([remap[int(s)] for s in set_id], dtype=np.int32),
             slot=slot, exact=exact, kind=np.array(kinds, dtype=np.int8))
    print(f"{len(exact)} integrals over {len(used)} shell sets; |exact| in "
          f"[{np.abs(exact).min():.2e}, {np.abs(exact).max():.2e}]")


if __name__ == "__main__":
    main()
