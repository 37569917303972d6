"""Autotuner for the per-class kernel choice of the HGP ERI engine.

For every quartet class it picks, on the GPU it runs on:
  * the kernel variant: generated register kernel (lane group per shell
    quartet) or the LDS wave kernel (1, 2 or 4 wavefronts per shell quartet), and
  * for the register kernel, the lanes per quartet G as a function of the
    contraction depth: G = pow2floor(K_bra * K_ket / tau), clamped to [1, 64],
    with tau chosen per class; for the wave kernel, the wavefront count.
Method: each candidate setting is applied to all classes at once, one
execute is profiled with HIP events between launches (median of ``reps``),
and each class keeps its fastest candidate.  The result is written as
hgp_tuning.txt rows "class bra_bucket ket_bucket variant G", read by mi_create.

Usage (on the GPU):  python -m mi_amd.autotune [--out PATH] [--reps N]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import configs
from .eri import Context
from .hgp_tables import quartet_classes

NBUCKET = 5
BUCKET_K = [1, 3, 7, 20, 48]          # representative pair contraction per bucket
TAUS = [1, 2, 4, 8, 16, 32, 64, 1 << 30]


def pow2floor(x):
    g = 1
    while g * 2 <= x and g < 64:
        g *= 2
    return g


def rows_for(cls_choice, has_reg, w_ok):
    """cls_choice[c] = (variant, tau) for the register kernel or (0, nw) for
    the wave kernel with nw wavefronts per quartet."""
    rows = []
    for c, (var, tau) in cls_choice.items():
        if var == 1 and not has_reg[c]:
            var, tau = 0, max([n for n in (1, 2, 4) if w_ok[c].get(n)] or [1])
        if var == 0 and not w_ok[c].get(int(tau), False):
            tau = max([n for n in (1, 2, 4) if w_ok[c].get(n)] or [1])
        for bb in range(NBUCKET):
            for kb in range(NBUCKET):
                G = pow2floor(max(1.0, BUCKET_K[bb] * BUCKET_K[kb] / tau)) if var == 1 else int(tau)
                rows.append((c, bb, kb, var, G))
    return rows


def class_times(ctx, fb, rows, reps):
    import torch
    ctx.set_tuning(rows)
    plan = ctx.plan(fb, order=1)
    out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream
    launches = plan.launches()
    prof = np.array([plan.profile(out.data_ptr(), sptr) for _ in range(reps)])
    med = np.median(prof, axis=0)
    per = {}
    for i, (cls, nqs, nqp, nblk, kern) in enumerate(launches):
        per[cls] = per.get(cls, 0.0) + float(med[i + 1])
    plan.close()
    return per


def autotune(workloads, reps=7, verbose=True):
    classes = quartet_classes()
    cid = {c: k for k, c in enumerate(classes)}
    ctx = Context(0)
    ctx.set_streams(1)
    # which variants each class has: probe the library by setting a row
    has_reg, w_ok = {}, {}
    for k in range(len(classes)):
        try:
            ctx.set_tuning([(k, 0, 0, 1, 1)])
            has_reg[k] = True
        except Exception:
            has_reg[k] = False
        w_ok[k] = {}
        for nw in (1, 2, 4):
            try:
                ctx.set_tuning([(k, 0, 0, 0, nw)])
                w_ok[k][nw] = True
            except Exception:
                w_ok[k][nw] = False
    cands = [(0, 1), (0, 2), (0, 4)] + [(1, t) for t in TAUS]
    total = {k: {} for k in range(len(classes))}
    for var, tau in cands:
        choice = {k: (var, tau) for k in range(len(classes))}
        rows = rows_for(choice, has_reg, w_ok)
        for fb in workloads:
            per = class_times(ctx, fb, rows, reps)
            for cls, t in per.items():
                k = cid[cls]
                if (var == 1 and not has_reg[k]) or (var == 0 and not w_ok[k].get(tau)):
                    continue
                total[k][(var, tau)] = total[k].get((var, tau), 0.0) + t
    best = {}
    for k in range(len(classes)):
        if not total[k]:
            best[k] = (1 if has_reg[k] else 0, 4)
            continue
        (var, tau), t = min(total[k].items(), key=lambda kv: kv[1])
        best[k] = (var, tau)
        if verbose:
            alt = ", ".join(f"{'W' + str(tt) if v == 0 else 'R/' + str(tt if tt < 1e6 else 'inf')}:{x * 1e3:.1f}"
                            for (v, tt), x in sorted(total[k].items(), key=lambda kv: kv[1])[:4])
            print(f"[autotune] {''.join('spdf'[x] for x in classes[k])}: "
                  f"{'W nw=' + str(tau) if var == 0 else 'R tau=' + str(tau)}  ({alt} us)", flush=True)
    ctx.close()
    return rows_for(best, has_reg, w_ok), best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "hgp_tuning.txt"))
    ap.add_argument("--reps", type=int, default=7)
    args = ap.parse_args()
    os.environ["MI_HGP_TUNING"] = "none"  # tune from the heuristic baseline
    wls = [configs.ethane_split_valence(0).flat()]
    rng_mols = configs.batch_molecules(1, nmol=2)
    wls += [m.flat() for m in rng_mols]
    rows, best = autotune(wls, args.reps)
    with open(args.out, "w") as f:
        f.write("# HGP ERI kernel choice per (class, bra bucket, ket bucket): variant (1 register kernel, "
                "0 LDS wave kernel) and lanes per quartet G.\n# Written by mi_amd/autotune.py on the GPU it "
                "describes; columns: class bra_bucket ket_bucket variant G\n")
        for r in rows:
            f.write(" ".join(str(int(x)) for x in r) + "\n")
    print(f"[autotune] wrote {args.out} ({len(rows)} rows)", flush=True)


if __name__ == "__main__":
    sys.exit(main())
