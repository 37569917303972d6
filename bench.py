#!/usr/bin/env python3
"""Throughput benchmark of the HGP two-electron integral path on one MI355X.

A step is one full pass of the hot path over the 60-function split-valence
ethane-like workload (seed 0): the shell-pair kernel plus every class launch
(issued as fused launch units), writing all 1,675,365 unique (ij|kl) in the
chrr order into a device-resident buffer.  Inputs (basis topology and
primitives) are resident in HBM before the timed region.

Prints ONE JSON line.  Extra keys:
  legs          the other workloads, each timed the same way (HIP events,
                inputs resident on the device):
                  batched_molecules   512 molecules of 3-24 atoms: plain
                                      per-molecule plans vs one batched plan
                                      replayed as a graph, FP64 and mixed
                  jk_cluster_40atoms  direct J/K, 920 functions with f shells:
                                      plain FP64 vs screened vs mixed+screened;
                                      its ERIs checked on 100,100 sampled shell
                                      quartets against the FP64 and 80-bit
                                      oracles (eri_spot_check)
                  serving             seeded Poisson schedule through the
                                      continuous-batching server: p50/p99
                                      latency and throughput, FP64 and mixed
  roofline      dominant kernel symbol: algorithmic FP64 FLOP per launch
                (FLOP model of mi_amd/flops.py) / its mean HIP-event duration
                between launch units, against the FP64 vector peak (the
                recurrences do not use MFMA); traffic from rocprofv3 FETCH/WRITE
                counters (MEASURED_TRAFFIC_BYTES).
Other modes: --calibrate-copy MIB (HBM counter calibration), --profile-batch
NMOL (one batched execute for counter profiles).
  cpu_baseline  the CPU oracle (C restatement of the reference HGP algorithm,
                OpenMP over shell quartets) on this host, whole workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "molecularintegrals.jl_amd"))

METRIC = "unique ERIs/s, 1 GPU (plain vs batched+mixed @1e-10 Eh); p50/p99 latency"
FP64_VECTOR_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of a kernel symbol on the default workload, from two
# rocprofv3 --pmc runs (FETCH_SIZE x 2.000 and WRITE_SIZE x 1.000, the factors
# a 256 MiB device copy gives); profiles/hbm_traffic_ethane_split_valence_60bf.md
# holds the commands and the per-kernel table.
MEASURED_TRAFFIC_BYTES = {"hgp_wf_kernel": 3855600, "hgp_rf2_kernel": 14963400, "hgp_rf0_kernel": 11218500,
                          "hgp_rf1_kernel": 11617300, "hgp_wgf_kernel": 2572000}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(fb, n_unique, budget_s=12.0):
    """Time the oracle (test infrastructure, used only as the reported CPU
    baseline) on this host: all threads of the CPU share, repeated runs of the
    whole workload until ~budget_s, median run time."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=threads)  # warm-up
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end and len(times) < 400:
        t = time.perf_counter()
        O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=threads)
        times.append(time.perf_counter() - t)
    t1 = []
    for _ in range(3):
        t = time.perf_counter()
        O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=1)
        t1.append(time.perf_counter() - t)
    med = float(np.median(times))
    return {"value": n_unique / med, "unit": "unique ERIs/s", "cores": threads, "kind": "port",
            "sample": f"whole 60-function workload, {len(times)} runs after 1 warm-up, median {med * 1e3:.2f} ms; "
                      f"oracle = C restatement of the reference HGP algorithm (D1-D3 corrected), OpenMP dynamic "
                      f"over shell quartets",
            "single_thread_value": n_unique / float(np.median(t1))}


def timed(fn, reps, stream):
    """Mean ms of fn() over reps calls, HIP events on stream (one warm-up)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def plan_flops(plan):
    """Algorithmic FLOPs of one execute of a plan (its class launches)."""
    from mi_amd import flops
    return sum(flops.class_flops(cls, nqs, nqp) for cls, nqs, nqp, _, _ in plan.launches())


def leg_batched(nmol, seed=0):
    """Many molecules of uneven size (3-24 atoms, split-valence recipe):
    the plain path (one plan per molecule, executed one after the other) vs
    one batched plan replayed as a captured graph, FP64 and the mixed form
    (FP32 VRR below tau = 1e-8, quartets with Schwarz bound product below
    1e-12 skipped -- both inside the 1e-10 error budget).  Errors: mixed vs
    FP64 over every integral (per molecule, on the device); FP64 vs the CPU
    oracle on the three smallest molecules."""
    import torch
    from mi_amd import Context, configs
    from oracle import oracle as O

    mols = configs.batch_molecules(seed, nmol=512)[:nmol]
    fbs = [m.flat() for m in mols]
    ctx = Context(0)
    st = torch.cuda.current_stream()
    sp = st.cuda_stream
    t = time.time()
    bplan = ctx.batch_plan(fbs)
    t_plan = time.time() - t
    n = bplan.n_unique
    off = bplan.mol_offsets()
    log(f"[bench] batched: {nmol} molecules, {n:.3e} unique ERIs ({n * 8 / 1e9:.1f} GB), plan {t_plan:.2f} s")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    bplan.capture(out.data_ptr())
    ms_fp64 = timed(lambda: bplan.replay(sp), 2, st)
    flops_fp64 = plan_flops(bplan)
    # plain: per-molecule plans, eager execute into the same slices
    plans = [ctx.plan(fb) for fb in fbs]
    base = out.data_ptr()
    def plain():
        for m, p in enumerate(plans):
            p.execute(base + 8 * int(off[m]), sp)
    ms_plain = timed(plain, 1, st)
    # CPU oracle spot check of the FP64 results
    small = sorted(range(nmol), key=lambda m: fbs[m].nbf)[:3]
    err_oracle = 0.0
    for m in small:
        ref = O.eri_unique_hgp(fbs[m], order=O.ORDER_CHRR, nthreads=8)
        got = out[int(off[m]):int(off[m + 1])].cpu().numpy()
        err_oracle = max(err_oracle, float(np.abs(got - ref).max()))
    # mixed
    mplan = ctx.batch_plan(fbs, precision=1, tau=1e-8, schwarz_cut=1e-12)
    mplan.capture(out.data_ptr())
    ms_mixed = timed(lambda: mplan.replay(sp), 2, st)
    flops_mixed = plan_flops(mplan)  # the quartets computed (screened ones never enter a task)
    scratch = torch.empty(int(max(off[m + 1] - off[m] for m in range(nmol))), dtype=torch.float64, device="cuda")
    err_mixed = 0.0
    for m, p in enumerate(plans):
        k = int(off[m + 1] - off[m])
        p.execute(scratch.data_ptr(), sp)
        err_mixed = max(err_mixed, float((out[int(off[m]):int(off[m]) + k] - scratch[:k]).abs().max()))
    for p in plans:
        p.close()
    bplan.close()
    mplan.close()
    del out, scratch
    torch.cuda.empty_cache()
    ctx.close()
    r = {"molecules": nmol, "unique_eris": int(n), "output_gb": n * 8 / 1e9,
         "plain_per_molecule": {"ms": ms_plain, "value": n / (ms_plain * 1e-3)},
         "batched_fp64_graph": {"ms": ms_fp64, "value": n / (ms_fp64 * 1e-3),
                                "output_write_gbs": n * 8 / (ms_fp64 * 1e-3) / 1e9,
                                "output_write_frac_hbm": n * 8 / (ms_fp64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "algorithmic_flop": flops_fp64,
                                "achieved_tflops": flops_fp64 / (ms_fp64 * 1e-3) / 1e12,
                                "frac_fp64_peak": flops_fp64 / (ms_fp64 * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS},
         "batched_mixed_graph": {"ms": ms_mixed, "value": n / (ms_mixed * 1e-3), "tau": 1e-8,
                                 "schwarz_cut": 1e-12, "algorithmic_flop_computed": flops_mixed,
                                 "achieved_tflops": flops_mixed / (ms_mixed * 1e-3) / 1e12},
         "speedup_batched_mixed_vs_plain": ms_plain / ms_mixed,
         "max_abs_err_mixed_vs_fp64": err_mixed,
         "max_abs_err_fp64_vs_cpu_oracle_3_smallest": err_oracle}
    log(f"[bench] batched: plain {ms_plain:.1f} ms, batched fp64 {ms_fp64:.1f} ms, mixed {ms_mixed:.1f} ms, "
        f"speedup {ms_plain / ms_mixed:.2f}x, err mixed {err_mixed:.2e}, err vs oracle {err_oracle:.2e}")
    log(f"[bench] batched: fp64 {flops_fp64 / (ms_fp64 * 1e-3) / 1e12:.2f} TFLOP/s algorithmic "
        f"({100 * flops_fp64 / (ms_fp64 * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS:.1f}% of FP64 peak), "
        f"mixed {flops_mixed / (ms_mixed * 1e-3) / 1e12:.2f} TFLOP/s on the quartets computed")
    return r


def eri_spot_check(ctx, fb, nsets, seed=0):
    """Checker for a basis whose unique array is too large to hold (the
    920-function cluster: 9e10 integrals): nsets random sets of 4 distinct
    shells, each a sub-basis with the cluster's own norms (its integrals are
    the cluster's), computed in one batched plan in plain FP64 and in the
    mixed, screened form, against the CPU oracle of each sub-basis.  Screening
    and FP32 choices depend only on the quartet's pairs, so the mixed values
    are those the cluster's own J/K build uses.  Returns max |dERI| per form."""
    import torch
    from mi_amd import sub_basis
    from oracle import oracle as O

    rng = np.random.Generator(np.random.PCG64(seed + 77))
    subs = [sub_basis(fb, rng.choice(fb.nshell, 4, replace=False)) for _ in range(nsets)]
    # the reference algorithm in FP64, and in 80-bit: with f shells the FP64
    # algorithm loses digits to HRR cancellation on some quartets (the
    # reference keeps a diffuse low-L shell on A), the 80-bit one does not
    refs = {"fp64_oracle": [O.eri_unique_hgp(sb, order=O.ORDER_CHRR, nthreads=16) for sb in subs],
            "80bit_oracle": [O.eri_unique_hgp_ext(sb, order=O.ORDER_CHRR, nthreads=16) for sb in subs]}
    sp = torch.cuda.current_stream().cuda_stream
    res = {"shell_sets": nsets, "shell_quartets": nsets * 55,
           "integrals": int(sum(r.size for r in refs["fp64_oracle"]))}
    for name, kw in (("fp64", {}), ("mixed_screened", dict(precision=1, tau=1e-8, schwarz_cut=1e-12))):
        plan = ctx.batch_plan(subs, **kw)
        out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr(), sp)
        got = out.cpu().numpy()
        off = plan.mol_offsets()
        for rn, rl in refs.items():
            d = [np.abs(got[off[m]:off[m + 1]] - rl[m]) for m in range(nsets)]
            res[f"max_abs_err_{name}_vs_{rn}"] = max(float(x.max()) for x in d)
            # integrals outside the parity bar (1e-12 absolute or 1e-10 relative), and above 1e-10 absolute
            res[f"n_outside_parity_bar_{name}_vs_{rn}"] = int(sum(
                np.count_nonzero((x > 1e-12) & (x > 1e-10 * np.abs(r))) for x, r in zip(d, rl)))
            res[f"n_above_1e-10_{name}_vs_{rn}"] = int(sum(np.count_nonzero(x > 1e-10) for x in d))
        plan.close()
    res["max_abs_err_fp64_oracle_vs_80bit_oracle"] = max(
        float(np.abs(a - b).max()) for a, b in zip(refs["fp64_oracle"], refs["80bit_oracle"]))
    res["n_above_1e-10_fp64_oracle_vs_80bit_oracle"] = int(sum(
        np.count_nonzero(np.abs(a - b) > 1e-10) for a, b in zip(refs["fp64_oracle"], refs["80bit_oracle"])))
    res["max_abs_eri"] = max(float(np.abs(r).max()) for r in refs["fp64_oracle"])
    return res


def leg_jk(seed=0):
    """Direct J/K on the 40-atom cluster (920 functions, l <= 3): plain FP64
    (no screening) vs screening at 1e-12 vs mixed (FP32 VRR below 1e-8) with
    screening; errors of J and K against the plain FP64 GPU result, and the
    cluster's ERIs checked against the CPU oracle on 1820 sampled sets of 4
    shells (100,100 shell quartets)."""
    import torch
    from mi_amd import Context, _lib, configs

    wl, D = configs.cluster_40(seed)
    fb = wl.flat()
    n = fb.nbf
    nu = n * (n + 1) * (n * n + n + 2) // 8
    ctx = Context(0)
    st = torch.cuda.current_stream()
    sp = st.cuda_stream
    dD = torch.from_numpy(D).cuda()
    J = torch.empty(n, n, dtype=torch.float64, device="cuda")
    K = torch.empty_like(J)
    res, mats = {}, {}
    for name, kw, reps in (("plain_fp64", {}, 1), ("fp64_screened", dict(schwarz_cut=1e-12), 2),
                           ("mixed_screened", dict(precision=1, tau=1e-8, schwarz_cut=1e-12), 2)):
        plan = ctx.plan(fb, order=_lib.MI_ORDER_JK, **kw)
        ms = timed(lambda: plan.jk_execute(dD.data_ptr(), J.data_ptr(), K.data_ptr(), sp), reps, st)
        mats[name] = (J.cpu().numpy(), K.cpu().numpy())
        res[name] = {"ms": ms, "value": nu / (ms * 1e-3)}
        plan.close()
        log(f"[bench] J/K {name}: {ms:.1f} ms, {nu / (ms * 1e-3):.3e} unique ERIs/s (never stored)")
    J0, K0 = mats["plain_fp64"]
    for name in ("fp64_screened", "mixed_screened"):
        res[name]["max_abs_err_J"] = float(np.abs(mats[name][0] - J0).max())
        res[name]["max_abs_err_K"] = float(np.abs(mats[name][1] - K0).max())
    # ERIs of the cluster against the CPU oracle on sampled shell quartets
    res["eri_spot_check"] = eri_spot_check(ctx, fb, 1820, seed)
    sc = res["eri_spot_check"]
    log(f"[bench] J/K cluster ERI spot check: {sc['shell_quartets']} shell quartets, {sc['integrals']} integrals; "
        f"max |dERI| vs 80-bit oracle: fp64 {sc['max_abs_err_fp64_vs_80bit_oracle']:.2e}, mixed+screened "
        f"{sc['max_abs_err_mixed_screened_vs_80bit_oracle']:.2e} (FP64 oracle vs 80-bit: "
        f"{sc['max_abs_err_fp64_oracle_vs_80bit_oracle']:.2e}); integrals above 1e-10: fp64 "
        f"{sc['n_above_1e-10_fp64_vs_80bit_oracle']}, mixed {sc['n_above_1e-10_mixed_screened_vs_80bit_oracle']}")
    ctx.close()
    res.update({"nbf": n, "unique_eris": nu, "screening": 1e-12,
                "speedup_mixed_screened_vs_plain": res["plain_fp64"]["ms"] / res["mixed_screened"]["ms"]})
    return res


def leg_serving(nreq, rate, seed=0):
    """Poisson request schedule (2-60 atoms, split-valence or cluster recipe
    50/50; unique ERIs up to 120 functions, J/K beyond) replayed in real time
    through the continuous-batching server; plain FP64 and mixed form."""
    from mi_amd import configs
    from mi_amd.serving import serve

    reqs = configs.serving_requests(seed=seed, nreq=nreq, rate_per_s=rate)
    out = {"requests": nreq, "rate_per_s": rate, "eri_requests": sum(r.kind == "eri" for r in reqs),
           "jk_requests": sum(r.kind == "jk" for r in reqs),
           "nbf_range": [min(r.workload.nbf for r in reqs), max(r.workload.nbf for r in reqs)]}
    res = {}
    for name, mixed in (("mixed", True), ("plain_fp64", False)):
        r = serve(reqs, mixed=mixed)
        res[name] = r
        lat = {k: np.array([r.latency_s[q.rid] for q in reqs if q.kind == k]) * 1e3 for k in ("eri", "jk")}
        out[name] = {"value": r.throughput, "unit": "unique ERIs/s", "p50_ms": r.p(0.5) * 1e3,
                     "p99_ms": r.p(0.99) * 1e3, "makespan_s": r.makespan_s, "batches": r.stats["batches"],
                     "peak_device_gb": r.stats["peak_bytes"] / 1e9,
                     "p50_ms_eri_requests": float(np.median(lat["eri"])) if lat["eri"].size else None,
                     "p50_ms_jk_requests": float(np.median(lat["jk"])) if lat["jk"].size else None}
        log(f"[bench] serving {name}: p50 {r.p(0.5) * 1e3:.1f} ms p99 {r.p(0.99) * 1e3:.1f} ms "
            f"{r.throughput:.3e} unique ERIs/s, makespan {r.makespan_s:.2f} s")
    err = {"eri": 0.0, "jk": 0.0}
    for q in reqs:
        a = [t.cpu().numpy() for t in res["plain_fp64"].outputs[q.rid]]
        b = [t.cpu().numpy() for t in res["mixed"].outputs[q.rid]]
        k = 0 if q.kind == "eri" else 1
        err[q.kind] = max(err[q.kind], max(float(np.abs(x - y).max()) for x, y in zip(a[k:], b[k:])))
    out["mixed"]["max_abs_err_eri_vs_fp64"] = err["eri"]
    out["mixed"]["max_abs_err_jk_vs_fp64"] = err["jk"]
    del res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="only the headline workload")
    ap.add_argument("--batch-mols", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--profile-batch", type=int, default=0, metavar="NMOL",
                    help="only execute one batched plan of NMOL molecules of 3-24 atoms (seed 1), one stream, "
                         "for counter profiles, and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8,
                          "copies": 5}), flush=True)
        return
    if args.profile_batch:
        import torch
        from mi_amd import Context, configs
        fbs = [m.flat() for m in configs.batch_molecules(1, nmol=args.profile_batch)]
        ctx = Context(0)
        ctx.set_streams(1)
        plan = ctx.batch_plan(fbs)
        out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        print(json.dumps({"profile_batch_molecules": args.profile_batch, "unique_eris": plan.n_unique,
                          "launches": [["".join("spdf"[x] for x in l[0]), l[4], l[1], l[2]]
                                       for l in plan.launches()]}), flush=True)
        plan.close()
        ctx.close()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)

    import torch
    from mi_amd import Context, configs, flops
    from mi_amd.hgp_tables import register_group

    wl = configs.ethane_split_valence(args.seed)
    fb = wl.flat()
    ctx = Context(0)
    plan = ctx.plan(fb, order=1)
    n = plan.n_unique
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    log(f"[bench] workload {wl.name} nbf={fb.nbf} unique={n} stats={plan.stats()}")

    for _ in range(args.warmup):
        plan.execute(out.data_ptr(), sptr)
    # stream lanes from measured launch times (plan set-up; results unchanged)
    plan.balance(out.data_ptr(), sptr)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        plan.execute(out.data_ptr(), sptr)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_per_step = e0.elapsed_time(e1) / args.steps
    value = n / (ms_per_step * 1e-3)
    log(f"[bench] {ms_per_step * 1e3:.1f} us/step, {value:.3e} unique ERIs/s")

    # Per-unit durations (HIP events between the launch units of an execute,
    # issued as executed, on the launch stream).  A unit is one class launch
    # or a fused launch of several classes (hgp_rf<g>_kernel: register
    # kernels of fusion group g; hgp_wf_kernel: wave kernels; hgp_wgf_kernel:
    # global-table kernels).  The roofline
    # is taken for the dominant kernel SYMBOL: hgp_r_<class> is one class,
    # the others serve several classes, so their time and FLOPs are summed
    # over the units of one step and the mean launch duration is
    # total / launches (what rocprofv3 --stats reports per symbol).
    launches = plan.launches()
    units = plan.units()
    nprof = max(5, min(args.steps, 50))
    acc = np.zeros(len(units) + 1)
    for _ in range(nprof):
        acc += np.array(plan.profile_units(out.data_ptr(), sptr))
    mean_ms = acc / nprof
    cname = lambda cls: "".join("spdf"[x] for x in cls)
    per = []
    step_flops = 0
    sym = {}
    for i, (kind, nw, members) in enumerate(units):
        f = sum(flops.class_flops(launches[m][0], launches[m][1], launches[m][2]) for m in members)
        step_flops += f
        kern = launches[members[0]][4]
        if kind == 1:
            name = f"hgp_rf{register_group(launches[members[0]][0])}_kernel"
        elif kind == 2:
            name = "hgp_wf_kernel"
        elif kind == 3:
            name = "hgp_wgf_kernel"
        else:
            name = {"R": "hgp_r_" + cname(launches[members[0]][0]), "W": "hgp_w_kernel",
                    "T": "hgp_w_kernel", "G": "hgp_wg_kernel", "X": "hgp_wx_kernel"}[kern[0]]
        label = " ".join(cname(launches[m][0]) for m in members) + f" [{name}{'' if kind else ' ' + kern}]"
        per.append((mean_ms[i + 1], label, f))
        t, fl, nl = sym.get(name, (0.0, 0, 0))
        sym[name] = (t + mean_ms[i + 1], fl + f, nl + 1)
    name, (dom_ms_total, dom_flops, dom_n) = max(sym.items(), key=lambda kv: kv[1][0])
    achieved = dom_flops / (dom_ms_total * 1e-3) / 1e12
    for ms_, label, f in sorted(per, key=lambda x: -x[0])[:8]:
        log(f"[bench]   {ms_ * 1e3:8.2f} us {f / (ms_ * 1e-3) / 1e12:7.3f} TFLOP/s  {label}")
    for nm, (t, fl, nl) in sorted(sym.items(), key=lambda kv: -kv[1][0])[:5]:
        log(f"[bench]   symbol {nm}: {nl} launch(es), {t * 1e3:.2f} us/step, {fl / (t * 1e-3) / 1e12:.3f} TFLOP/s")
    log(f"[bench]   pair kernel {mean_ms[0] * 1e3:.2f} us; sum of units {mean_ms.sum() * 1e3:.1f} us")

    # parity spot check of the timed output against the committed fixture
    g = np.load(os.path.join(ROOT, "tests", "golden", "ethane_split_valence_seed0.npz"))
    got = out.cpu().numpy()
    err = float(np.max(np.abs(got[g["slots"]] - g["values"]))) if args.seed == 0 else None

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "unique ERIs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded basis and geometry)",
        "config": {"workload": "ethane-like 8 atoms, split-valence+polarisation, 60 Cartesian functions "
                               f"(seed {args.seed}), plain FP64 HGP, all unique ERIs in chrr order",
                   "nbf": fb.nbf, "unique_eris": n, "shell_quartets": plan.stats()["shell_quartets"],
                   "primitive_quartets": plan.stats()["primitive_quartets"], "precision": "fp64",
                   "order": "chrr"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
                     "traffic": MEASURED_TRAFFIC_BYTES.get(name) if args.seed == 0 else None,
                     "kernel": name, "kernel_launches_per_step": dom_n,
                     "kernel_mean_us": dom_ms_total * 1e3 / dom_n, "kernel_algorithmic_flop": dom_flops / dom_n,
                     "step_algorithmic_flop": step_flops,
                     "step_achieved_tflops": step_flops / (ms_per_step * 1e-3) / 1e12,
                     "step_output_write_gbs": n * 8 / (ms_per_step * 1e-3) / 1e9,
                     "step_output_write_frac_hbm": n * 8 / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "max_abs_err_vs_golden": err,
    }
    if not args.no_cpu:
        log("[bench] timing CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(fb, n, args.cpu_budget_s)
    plan.close()
    del out
    torch.cuda.empty_cache()
    if not args.no_legs:
        legs = {}
        for name, fn in (("batched_molecules", lambda: leg_batched(args.batch_mols)),
                         ("jk_cluster_40atoms", leg_jk),
                         ("serving", lambda: leg_serving(args.serve_requests, args.serve_rate))):
            try:
                legs[name] = fn()
            except Exception as e:  # report, keep the headline line
                legs[name] = {"error": f"{type(e).__name__}: {e}"}
                log(f"[bench] leg {name} failed: {e}")
            torch.cuda.empty_cache()
        result["legs"] = legs
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
