"""Autotuner for the per-class kernel choice of the HGP ERI engine.

For every quartet class it picks, on the GPU it runs on:
  * the kernel variant: generated register kernel (lane group per shell
    quartet), the LDS wave kernel (1, 2 or 4 wavefronts per shell quartet;
    op tables staged in LDS or read from global memory) or the global-table
    kernel (1, 2, 4 or 8 wavefronts per block), and
  * for the register kernel, the lanes per quartet G as a function of the
    contraction depth: G = pow2floor(K_bra * K_ket / tau), clamped to [1, 64],
    with tau chosen per class; for the wave kernels, the wavefront count.
Three regimes are tuned apart: small class launches (fewer than 65536 shell
quartets, latency-bound: more lanes per quartet pay), large ones
(throughput-bound) -- a plan picks one of these per class from its quartet
count -- and the launches of Coulomb/exchange plans (MI_ORDER_JK: the block
is contracted with the density instead of written out).
Method: each candidate setting is applied to all classes at once, one
execute per workload is profiled with HIP events between launches (median
of ``reps``), and each class keeps, per regime, the candidate with the
smallest sum over that regime's workloads of its time relative to the
workload's best candidate.  The result is written as hgp_tuning.txt rows
"class bra_bucket ket_bucket variant G regime", read by mi_create.

Usage (on the GPU):  python -m mi_amd.autotune [--out PATH] [--reps N] [--regimes 0,1,2 [--merge]]
(--merge keeps the rows of the regimes not tuned from the existing file);
python -m mi_amd.autotune --refine 0 (then --refine 1) refines the small-
(large-) launch regime on whole-execute times (refine_regime).
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import _lib, configs
from .eri import Context
from .hgp_tables import quartet_classes

NBUCKET = 5
BUCKET_K = [1, 3, 7, 20, 48]          # representative pair contraction per bucket
TAUS = [1, 2, 4, 8, 16, 32, 64, 1 << 30]
# kernel choices: (0, nw) LDS wave kernel, (2, nw) global-table kernel, (3, nw)
# LDS wave kernel with global op tables, (1, tau) register kernel, (5, nw) Rys
# quadrature kernel (the same integrals by Rys quadrature; HGP transfer)
RYS_CANDS = [(5, 1), (5, 2), (5, 4), (6, 16), (6, 32), (6, 64)]
CANDS = [(0, 1), (0, 2), (0, 4), (2, 1), (2, 2), (2, 4), (2, 8), (3, 1), (3, 2), (3, 4)] + [(1, t) for t in TAUS]


def pow2floor(x):
    g = 1
    while g * 2 <= x and g < 64:
        g *= 2
    return g


def fallback(c, w_ok, g_ok):
    """A wave-kernel setting that exists for class c."""
    w = [n for n in (1, 2, 4) if w_ok[c].get(n)]
    return (0, max(w)) if w else (2, 4)


def available(c, var, tau, has_reg, w_ok, g_ok, t_ok=None):
    """Whether class c has kernel choice (var, tau); t_ok carries the Rys
    kernel's wavefront counts under key 5."""
    if var == 1:
        return has_reg[c]
    if var in (5, 6):
        return t_ok is not None and t_ok[c].get((var, int(tau)), False)
    ok = {0: w_ok, 2: g_ok, 3: t_ok}.get(var)
    return ok is not None and ok[c].get(int(tau), False)


def rows_for(cls_choice, has_reg, w_ok, g_ok, t_ok=None):
    """cls_choice[c] = (1, tau) for the register kernel, (0, nw) for the LDS
    wave kernel, (2, nw) for the global-table kernel, (3, nw) for the LDS wave
    kernel with global op tables."""
    rows = []
    for c, (var, tau) in cls_choice.items():
        if not available(c, var, tau, has_reg, w_ok, g_ok, t_ok):
            var, tau = fallback(c, w_ok, g_ok)
        for bb in range(NBUCKET):
            for kb in range(NBUCKET):
                G = pow2floor(max(1.0, BUCKET_K[bb] * BUCKET_K[kb] / tau)) if var == 1 else int(tau)
                rows.append((c, bb, kb, var, G))
    return rows


class JKWorkload:
    """A Coulomb/exchange build (random symmetric density) under plan options kw."""

    def __init__(self, fb, seed=0, **kw):
        self.fb, self.kw = fb, kw
        rng = np.random.Generator(np.random.PCG64(seed))
        d = rng.normal(size=(fb.nbf, fb.nbf)) / fb.nbf
        self.D = 0.5 * (d + d.T)


def class_times(ctx, wl, rows, reps):
    """Per-class time (ms) of one execute of workload wl: a FlatBasis, a list
    of them (a batched plan) or a JKWorkload."""
    import torch
    ctx.set_tuning(rows)
    sptr = torch.cuda.current_stream().cuda_stream
    if isinstance(wl, JKWorkload):
        plan = ctx.plan(wl.fb, order=_lib.MI_ORDER_JK, **wl.kw)
        D = torch.from_numpy(wl.D).cuda()
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        launches = plan.launches()
        prof = np.array([plan.jk_profile(D.data_ptr(), J.data_ptr(), K.data_ptr(), sptr) for _ in range(reps)])
    else:
        plan = ctx.batch_plan(wl) if isinstance(wl, list) else ctx.plan(wl, order=1)
        out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
        launches = plan.launches()
        prof = np.array([plan.profile(out.data_ptr(), sptr) for _ in range(reps)])
    med = np.median(prof, axis=0)
    per = {}
    for i, (cls, nqs, nqp, nblk, kern) in enumerate(launches):
        per[cls] = per.get(cls, 0.0) + float(med[i + 1])
    plan.close()
    return per


def autotune(workloads, regimes, reps=7, verbose=True):
    """workloads: FlatBasis, lists of them (batched plans) or JKWorkload;
    regimes[i]: the tuning regime (0 small class launches, 1 large, 2 J/K
    plans) workload i stands for.  Returns rows for the regimes present."""
    classes = quartet_classes()
    cid = {c: k for k, c in enumerate(classes)}
    ctx = Context(0)
    ctx.set_streams(1)
    # which variants each class has: probe the library by setting a row
    has_reg, w_ok, g_ok, t_ok = probe_variants(ctx, len(classes))
    cands = CANDS
    # times[w][k][cand]
    times = [{k: {} for k in range(len(classes))} for _ in workloads]
    for var, tau in cands:
        choice = {k: (var, tau) for k in range(len(classes))}
        rows = rows_for(choice, has_reg, w_ok, g_ok, t_ok)
        for wi, wl in enumerate(workloads):
            per = class_times(ctx, wl, rows, reps)
            for cls, t in per.items():
                k = cid[cls]
                if not available(k, var, tau, has_reg, w_ok, g_ok, t_ok):
                    continue
                times[wi][k][(var, tau)] = t
    def pick(wis):
        """Best candidate per class over the workloads wis (score: sum of
        time / best time per workload; a candidate must have run on every
        workload that has the class)."""
        best, score = {}, {}
        for k in range(len(classes)):
            tot = {}
            have = [wi for wi in wis if times[wi][k]]
            for wi in have:
                t0 = min(times[wi][k].values())
                for cand, t in times[wi][k].items():
                    tot[cand] = tot.get(cand, 0.0) + t / t0
            tot = {c: v for c, v in tot.items() if all(c in times[wi][k] for wi in have)}
            if tot:
                best[k] = min(tot.items(), key=lambda kv: kv[1])[0]
                score[k] = tot
        return best, score

    name = lambda v, tt: ({0: "W", 2: "G", 3: "T"}.get(v, "R/") + str(tt if tt < 1e6 else "inf"))
    rows = []
    tuned = sorted(set(regimes))
    picks = {regime: pick([wi for wi, r in enumerate(regimes) if r == regime]) for regime in tuned}
    for regime in tuned:
        best, score = picks[regime]
        others = [picks[r][0] for r in tuned if r != regime]
        choice = {}
        for k in range(len(classes)):
            if k in best:
                choice[k] = best[k]
            elif regime == 2:
                continue  # J/K plans fall back to the size regimes for classes not seen here
            elif any(k in o for o in others):  # class absent from this regime's workloads
                choice[k] = next(o[k] for o in others if k in o)
            else:
                choice[k] = (1, 4) if has_reg[k] else fallback(k, w_ok, g_ok)
            if verbose and k in score:
                alt = ", ".join(f"{name(v, tt)}:{x:.2f}"
                                for (v, tt), x in sorted(score[k].items(), key=lambda kv: kv[1])[:4])
                print(f"[autotune] regime {regime} {''.join('spdf'[x] for x in classes[k])}: "
                      f"{name(*choice[k])}  ({alt})", flush=True)
        rows += [r + (regime,) for r in rows_for(choice, has_reg, w_ok, g_ok, t_ok)]
    ctx.close()
    return rows


def read_rows(path):
    """Rows (class, bra_bucket, ket_bucket, variant, G, regime) of a tuning file."""
    rows = []
    for line in open(path):
        f = line.split()
        if not line.startswith("#") and len(f) == 6:
            rows.append(tuple(int(x) for x in f))
    return rows


def probe_variants(ctx, nclass):
    """Which kernel variants each class has (probed by setting a row)."""
    has_reg, w_ok, g_ok, t_ok = {}, {}, {}, {}
    for k in range(nclass):
        try:
            ctx.set_tuning([(k, 0, 0, 1, 1)])
            has_reg[k] = True
        except Exception:
            has_reg[k] = False
        w_ok[k], g_ok[k], t_ok[k] = {}, {}, {}
        for nw in (1, 2, 4, 8):
            for var, ok in ((0, w_ok[k]), (2, g_ok[k]), (3, t_ok[k])):
                try:
                    ctx.set_tuning([(k, 0, 0, var, nw)])
                    ok[nw] = True
                except Exception:
                    ok[nw] = False
        for var, gs in ((5, (1, 2, 4)), (6, (16, 32, 64))):  # Rys kernels
            for g in gs:
                try:
                    ctx.set_tuning([(k, 0, 0, var, g)])
                    t_ok[k][(var, g)] = True
                except Exception:
                    t_ok[k][(var, g)] = False
    return has_reg, w_ok, g_ok, t_ok


def step_time(ctx, wl, reps, n=40):
    """Median over reps of the mean time (ms) of n executes of workload wl:
    (FlatBasis or list of them, plan options) for the unique array, as
    executed (fused units, stream lanes balanced on measured times), or a
    JKWorkload for Coulomb/exchange builds."""
    import torch
    if isinstance(wl, JKWorkload):
        s = torch.cuda.current_stream()
        sp = s.cuda_stream
        plan = ctx.plan(wl.fb, order=_lib.MI_ORDER_JK, **wl.kw)
        D = torch.from_numpy(wl.D).cuda()
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        plan.jk_execute(D.data_ptr(), J.data_ptr(), K.data_ptr(), sp)
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(n):
                plan.jk_execute(D.data_ptr(), J.data_ptr(), K.data_ptr(), sp)
            e1.record(s)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / n)
        plan.close()
        return float(np.median(ts))
    fbs, kw = wl
    s = torch.cuda.current_stream()
    sp = s.cuda_stream
    plan = ctx.batch_plan(fbs, **kw) if isinstance(fbs, list) else ctx.plan(fbs, order=1, **kw)
    out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
    for _ in range(2 if n < 10 else 5):
        plan.execute(out.data_ptr(), sp)
    plan.balance(out.data_ptr(), sp)
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(n):
            plan.execute(out.data_ptr(), sp)
        e1.record(s)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) / n)
    plan.close()
    del out
    return float(np.median(ts))


def refine_regime(workloads, rows, regime=0, reps=5, n=40, passes=2, gain=0.015, verbose=True, checkpoint=None,
                  cands=None, min_l=0):
    """Coordinate descent on one tuning regime over whole-execute times.
    Per-class launch times, measured one class at a time, miss that classes
    share the chip (fused small launches; stream lanes), and that the
    quartets a screened or mixed plan keeps have another contraction profile
    than the full set.  So each class of the regime present in the workloads
    (heaviest first) tries every kernel choice and keeps the one that lowers
    the sum over workloads of execute time / starting time by more than
    ``gain``.  workloads: (FlatBasis or list of them, plan options).  Returns
    the updated rows (other regimes unchanged).  cands: the kernel choices
    tried (default CANDS); min_l: only classes of total angular momentum at
    least min_l."""
    classes = quartet_classes()
    ctx = Context(0)
    has_reg, w_ok, g_ok, t_ok = probe_variants(ctx, len(classes))
    cands = CANDS if cands is None else cands

    def valid(k, cand):
        return available(k, cand[0], cand[1], has_reg, w_ok, g_ok, t_ok)

    def with_choice(rs, k, cand):
        keep = [r for r in rs if not (r[0] == k and r[5] == regime)]
        return keep + [r + (regime,) for r in rows_for({k: cand}, has_reg, w_ok, g_ok, t_ok)]

    def times(rs):
        ctx.set_tuning(rs)
        return [step_time(ctx, wl, reps, n) for wl in workloads]

    base = times(rows)
    score = lambda t: sum(x / b for x, b in zip(t, base))
    cur = score(base)
    # classes of the regime's launches, heaviest first (per-launch times)
    import torch
    weight = {}
    ctx.set_tuning(rows)
    sp = torch.cuda.current_stream().cuda_stream
    for wl in workloads:
        if isinstance(wl, JKWorkload):
            plan = ctx.plan(wl.fb, order=_lib.MI_ORDER_JK, **wl.kw)
            D = torch.from_numpy(wl.D).cuda()
            J = torch.empty_like(D)
            K = torch.empty_like(D)
            prof = plan.jk_profile(D.data_ptr(), J.data_ptr(), K.data_ptr(), sp)
        else:
            fbs, kw = wl
            plan = ctx.batch_plan(fbs, **kw) if isinstance(fbs, list) else ctx.plan(fbs, order=1, **kw)
            out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
            prof = plan.profile(out.data_ptr(), sp)
        for i, (cls, nqs, _, _, _) in enumerate(plan.launches()):
            # regime 2: every launch of a J/K plan; else small / large launches
            if (regime == 2 or (nqs < 65536) == (regime == 0)) and sum(cls) >= min_l:
                k = classes.index(cls)
                weight[k] = weight.get(k, 0.0) + prof[i + 1]
        plan.close()
    order = sorted(weight, key=lambda k: -weight[k])
    name = lambda v, tt: ({0: "W", 2: "G", 3: "T", 5: "Y", 6: "Z"}.get(v, "R/") + str(tt if tt < 1e6 else "inf"))
    if verbose:
        print(f"[refine] regime {regime} start: " + ", ".join(f"{b:.3f} ms" for b in base) +
              f"; {len(order)} classes", flush=True)
    for p in range(passes):
        changed = False
        for k in order:
            best, bcand = cur, None
            for cand in cands:
                if not valid(k, cand):
                    continue
                sc = score(times(with_choice(rows, k, cand)))
                if sc < best * (1.0 - gain):
                    best, bcand = sc, cand
            if bcand is not None:
                rows = with_choice(rows, k, bcand)
                cur = best
                changed = True
                if checkpoint:
                    checkpoint(rows)
            if verbose:
                print(f"[refine] pass {p} {''.join('spdf'[x] for x in classes[k])}: "
                      f"{name(*bcand) if bcand else 'kept'}, score {cur:.3f}", flush=True)
        if not changed:
            break
    if verbose:
        print(f"[refine] end: " + ", ".join(f"{t:.3f} ms" for t in times(rows)), flush=True)
    ctx.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "hgp_tuning.txt"))
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--regimes", default="0,1,2", help="regimes to tune (comma separated)")
    ap.add_argument("--merge", action="store_true", help="keep the other regimes' rows of the existing --out file")
    ap.add_argument("--refine", type=int, default=-1, metavar="REGIME",
                    help="refine regime 0 (small launches, fused), 1 (large launches) or 2 (J/K plans) of the "
                         "existing --out file on whole-execute times")
    ap.add_argument("--passes", type=int, default=2, help="refinement passes over the classes")
    ap.add_argument("--rys-only", action="store_true",
                    help="with --refine: try only the Rys kernel choices (classes with L >= --min-l)")
    ap.add_argument("--min-l", type=int, default=0)
    args = ap.parse_args()
    if args.refine >= 0:
        rng = np.random.Generator(np.random.PCG64(9))
        small = configs.random_molecule(rng, 4, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat()
        batch = [m.flat() for m in configs.batch_molecules(1, nmol=48)]
        cluster = configs.random_molecule(rng, 16, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat()
        mixed = dict(precision=_lib.MI_PREC_MIXED, tau=1e-8, schwarz_cut=1e-12)
        if args.refine == 0:
            # small class launches (fused): the 60-function molecule and a
            # 4-atom cluster-recipe molecule
            wls, n = [(configs.ethane_split_valence(0).flat(), {}), (small, {})], 40
        elif args.refine == 2:
            # J/K plans: the mixed, screened J/K workloads of the per-class tuning
            jk16 = configs.random_molecule(rng, 16, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat()
            jk40 = configs.random_molecule(rng, 40, configs.SPLIT_HEAVY, configs.SPLIT_LIGHT).flat()
            wls, n = [JKWorkload(jk16, 1, **mixed), JKWorkload(jk40, 2, **mixed)], 2
        else:
            # large ones: 48 batched molecules in FP64 and in the mixed,
            # screened form, and a 16-atom cluster-recipe molecule (mixed)
            wls, n = [(batch, {}), (batch, mixed), (cluster, mixed)], 3
        rows = refine_regime(wls, read_rows(args.out), regime=args.refine, reps=args.reps, n=n,
                             passes=args.passes, checkpoint=lambda r: write_rows(args.out, r),
                             cands=RYS_CANDS if args.rys_only else None, min_l=args.min_l)
        write_rows(args.out, rows)
        return 0
    want = sorted({int(x) for x in args.regimes.split(",")})
    kept = []
    if args.merge and os.path.exists(args.out):
        for line in open(args.out):
            f = line.split()
            if line.startswith("#") or len(f) != 6:
                continue
            if int(f[5]) not in want:
                kept.append(tuple(int(x) for x in f))
    os.environ["MI_HGP_TUNING"] = "none"  # tune from the heuristic baseline
    # small class launches (latency-bound): the 60-function molecule and a
    # 4-atom cluster-recipe molecule with f shells; large ones (throughput): a
    # batch of 48 molecules of 3-24 atoms and a 16-atom cluster-recipe molecule
    # J/K plans: mixed precision with screening on a 16-atom cluster-recipe
    # molecule and a 40-atom split-valence one (the serving J/K requests)
    rng = np.random.Generator(np.random.PCG64(9))
    wls = [configs.ethane_split_valence(0).flat(),
           configs.random_molecule(rng, 4, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat(),
           [m.flat() for m in configs.batch_molecules(1, nmol=48)],
           configs.random_molecule(rng, 16, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat()]
    regimes = [0, 0, 1, 1]
    mixed = dict(precision=_lib.MI_PREC_MIXED, tau=1e-8, schwarz_cut=1e-12)
    wls += [JKWorkload(configs.random_molecule(rng, 16, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat(), 1, **mixed),
            JKWorkload(configs.random_molecule(rng, 40, configs.SPLIT_HEAVY, configs.SPLIT_LIGHT).flat(), 2, **mixed)]
    regimes += [2, 2]
    sel = [i for i, r in enumerate(regimes) if r in want]
    rows = kept + autotune([wls[i] for i in sel], [regimes[i] for i in sel], args.reps)
    write_rows(args.out, rows)


def write_rows(path, rows):
    with open(path, "w") as f:
        f.write("# HGP ERI kernel choice per (class, bra bucket, ket bucket, regime): variant (1 register "
                "kernel, 0 LDS wave kernel, 2 global-table kernel, 3 wave kernel with global tables, 5 Rys "
                "kernel, 6 Rys lane-group kernel) and G (lanes per quartet, or wavefronts);\n"
                "# regime 0 = class launches of fewer than 65536 shell quartets, 1 = larger ones, "
                "2 = launches of J/K plans.\n"
                "# Written by mi_amd/autotune.py on the GPU it describes; columns: class bra_bucket ket_bucket "
                "variant G regime\n")
        for r in rows:
            f.write(" ".join(str(int(x)) for x in r) + "\n")
    print(f"[autotune] wrote {path} ({len(rows)} rows)", flush=True)


if __name__ == "__main__":
    sys.exit(main())
