#!/usr/bin/env python3
"""Throughput benchmark of the HGP two-electron integral path on one MI355X.

Headline (``value``): the batched, mixed-precision form on 512 molecules of
3-24 atoms (split-valence recipe with d shells, seed 0; 2.43e10 unique ERIs,
194 GB of output resident in HBM).  A step is one replay of the captured
graph of one batched plan: the shell-pair kernel, the Schwarz-screened class
launches (quartets with Q_ab Q_cd < 1e-12 skipped, FP32 VRR where the
VRR-output bound product is below 1e-8), writing every unique (ij|kl) of
every molecule in the chrr order.  value = unique ERIs per second.  Inputs
(basis topology, primitives) are resident on the device before the clock.

The same line carries:
  plain          the plain GPU path on the same 512 molecules: one FP64 plan
                 per molecule, executed one after the other (the reference's
                 one-Basis-per-call driver on the GPU); speedup_vs_plain =
                 value / plain value
  batched_fp64   the batched plan in plain FP64 (graph replay)
  errors         every mixed integral against the FP64 per-molecule output
                 (max |d|, on the device); 10,240 slots sampled over all 512
                 molecules against the exact value of the reference algorithm
                 (binary128 oracle): mixed and plain FP64, max |d| and the
                 count outside the 1e-12 abs / 1e-10 rel bar
  roofline       dominant kernel symbol of the mixed execute: algorithmic
                 FP64 FLOP (FLOP model of mi_amd/flops.py) / its summed HIP
                 event time per launch (one launch per class, caller stream),
                 against the FP64 vector peak; traffic from rocprofv3
                 FETCH_SIZE/WRITE_SIZE of the same launch
                 (MEASURED_TRAFFIC_BYTES, profiles/)
  cpu_baseline   the CPU oracle (C restatement of the reference HGP algorithm,
                 OpenMP over shell quartets) on a fixed leading subset of the
                 same molecules, extrapolated as a rate
  legs           ethane_split_valence_60bf (the 60-function plain-FP64 step),
                 reference_timing_workloads (ethane STO-3G / 6-31G / cc-pVDZ,
                 HGP and Rys quadrature, beside the reference's published
                 single-thread times), jk_cluster_40atoms (direct J/K, 920
                 functions with f shells, with a 100,100-quartet ERI spot check
                 against the exact values), serving (Poisson requests of 2-60
                 atoms through the continuous-batching server, p50/p99) and
                 served_batch (the 512 molecules as ERI requests through the
                 server).
Other modes: --calibrate-copy MIB (HBM counter calibration), --profile-headline
(one execute of the headline plan on one stream, for rocprofv3),
--profile-batch NMOL (one FP64 batched execute, for counter profiles).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "molecularintegrals.jl_amd"))

METRIC = "unique ERIs/s, 1 GPU (plain vs batched+mixed @1e-10 Eh); p50/p99 latency"
FP64_VECTOR_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
MIXED = dict(precision=1, tau=1e-8, schwarz_cut=1e-12)
# HBM bytes per launch of a kernel symbol, from two rocprofv3 --pmc runs of
# `bench.py --profile-headline` (FETCH_SIZE x 2.000 and WRITE_SIZE x 1.000,
# the factors a 256 MiB device copy gives): profiles/hbm_traffic_batch512_mixed.md.
MEASURED_TRAFFIC_BYTES = {"hgp_ryg_kernel": 4046039491, "hgp_r_ppps": 18689379887, "hgp_r_dpps": 18179650698,
                          "hgp_r_dpds": 15032069152, "hgp_r_psps": 13542996320, "hgp_r_dsps": 12086832695,
                          "hgp_r_psss": 11161218090, "hgp_r_dspp": 9326708062, "hgp_r_pppp": 8431444566,
                          "hgp_r_ppss": 7119721388, "hgp_r_dpss": 6133995468, "hgp_r_ddps": 5588570871,
                          "hgp_r_dsss": 5131085530, "hgp_r_dsds": 3098223339, "hgp_wg_kernel": 2762015519,
                          "hgp_r_ssss": 2655462367, "hgp_r_ddss": 2273834727}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def omp_threads():
    """The OpenMP threads of the CPU baseline: OMP_NUM_THREADS (set to the
    host-core share of one GPU on the benchmark machines), else every core."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def cpu_baseline_batch(fbs, budget_s=15.0):
    """The oracle (test infrastructure, used only as the reported CPU baseline)
    on the leading molecules of the headline set, in order, until ~budget_s;
    the rate over that subset stands for the whole set (sizes are drawn
    independently per molecule)."""
    from oracle import oracle as O

    threads = omp_threads()
    O.eri_unique_hgp(fbs[0], order=O.ORDER_CHRR, nthreads=threads)  # warm-up
    done, n, t = 0, 0, 0.0
    while done < len(fbs) and t < budget_s:
        fb = fbs[done]
        t0 = time.perf_counter()
        O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=threads)
        t += time.perf_counter() - t0
        n += fb.nbf * (fb.nbf + 1) * (fb.nbf ** 2 + fb.nbf + 2) // 8
        done += 1
    total = sum(fb.nbf * (fb.nbf + 1) * (fb.nbf ** 2 + fb.nbf + 2) // 8 for fb in fbs)
    return {"value": n / t, "unit": "unique ERIs/s", "cores": threads, "kind": "port",
            "sample": f"molecules 0-{done - 1} of the 512 ({n:.3e} of {total:.3e} unique ERIs, "
                      f"{100 * n / total:.1f}%), whole arrays in FP64, {t:.1f} s; rate extrapolated to the set "
                      f"(CPU time for all 512: ~{total / (n / t):.0f} s); oracle = C restatement of the reference "
                      f"HGP algorithm (D1-D3, D5 corrected), OpenMP dynamic over shell quartets",
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def cpu_baseline_single(fb, n_unique, budget_s=10.0):
    from oracle import oracle as O

    threads = omp_threads()
    O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=threads)
    times = []
    t_end = time.time() + budget_s
    while time.time() < t_end and len(times) < 400:
        t = time.perf_counter()
        O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=threads)
        times.append(time.perf_counter() - t)
    med = float(np.median(times))
    # one thread as well: median of 5 runs after the warm-up above
    t1 = []
    for _ in range(5):
        t = time.perf_counter()
        O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=1)
        t1.append(time.perf_counter() - t)
    med1 = float(np.median(t1))
    return {"value": n_unique / med, "unit": "unique ERIs/s", "cores": threads, "kind": "port",
            "sample": f"whole 60-function array, {len(times)} runs, median {med * 1e3:.2f} ms",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "one_thread": {"value": n_unique / med1, "ms": med1 * 1e3, "sample": "median of 5 runs"}}


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


def symbol_of(launch):
    cname = "".join("spdf"[x] for x in launch[0])
    return {"R": "hgp_r_" + cname, "W": "hgp_w_kernel", "T": "hgp_w_kernel", "G": "hgp_wg_kernel",
            "X": "hgp_wx_kernel", "Y": "hgp_ry_kernel", "Z": "hgp_ryg_kernel"}[launch[4][0]]


def roofline_from_profile(plan, ms, step_ms, n_unique):
    """Dominant kernel symbol over one execute's class launches (per-launch
    HIP events on the caller's stream, ms[0] = shell-pair kernel): its
    algorithmic FLOP per launch over its mean launch time."""
    from mi_amd import flops
    sym = {}
    for l, t in zip(plan.launches(), ms[1:]):
        f = flops.class_flops(l[0], l[1], l[2])
        s = symbol_of(l)
        T, F, N = sym.get(s, (0.0, 0, 0))
        sym[s] = (T + t, F + f, N + 1)
    name, (t, f, nl) = max(sym.items(), key=lambda kv: kv[1][0])
    achieved = f / (t * 1e-3) / 1e12
    step_flops = plan_flops(plan)
    top = sorted(sym.items(), key=lambda kv: -kv[1][0])[:6]
    for nm, (tt, ff, nn) in top:
        log(f"[bench]   symbol {nm}: {nn} launch(es), {tt:.2f} ms/execute, {ff / (tt * 1e-3) / 1e12:.3f} TFLOP/s")
    return {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": MEASURED_TRAFFIC_BYTES.get(name),
            "kernel": name, "kernel_launches_per_step": nl, "kernel_mean_us": t * 1e3 / nl,
            "kernel_algorithmic_flop": f / nl, "step_algorithmic_flop": step_flops,
            "step_achieved_tflops": step_flops / (step_ms * 1e-3) / 1e12,
            "step_output_write_gbs": n_unique * 8 / (step_ms * 1e-3) / 1e9,
            "step_output_write_frac_hbm": n_unique * 8 / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "symbols_ms_per_execute": {nm: round(tt, 3) for nm, (tt, _, _) in top}}


def prim_quartets_unpruned(fbs):
    """Primitive quartets of every canonical shell quartet without the
    primitive-pair prune: per molecule sum_{p<=q} K_p K_q over shell pairs."""
    tot = 0
    for fb in fbs:
        k = np.diff(fb.shell_prim_off).astype(np.int64)
        kp = np.outer(k, k)[np.tril_indices(len(k))]
        s = int(kp.sum())
        tot += (s * s + int((kp * kp).sum())) // 2
    return tot


def sampled_exact_errors(fbs, off, got_fn, per_mol=20, seed=5):
    """per_mol random slots of every molecule against the binary128 oracle
    (exact value of the reference algorithm, Boys function exact for all T)."""
    from oracle import oracle as O
    rng = np.random.Generator(np.random.PCG64(seed))
    worst, nbad, nsl = 0.0, 0, 0
    for m, fb in enumerate(fbs):
        n = int(off[m + 1] - off[m])
        sl = np.unique(rng.integers(0, n, size=per_mol))
        ex = O.eri_slots_quad(fb, sl, order=O.ORDER_CHRR, accurate=True, nthreads=omp_threads())
        g = got_fn(int(off[m]) + sl)
        d = np.abs(g - ex)
        worst = max(worst, float(d.max()))
        nbad += int(np.count_nonzero((d > 1e-12) & (d > 1e-10 * np.abs(ex))))
        nsl += sl.size
    return {"slots": nsl, "max_abs_err_vs_exact": worst, "n_outside_parity_bar": nbad}


def headline(args):
    """512 molecules: the batched mixed graph (value), the plain per-molecule
    path, the batched FP64 graph, errors and the roofline."""
    import torch
    from mi_amd import Context, configs

    mols = configs.batch_molecules(args.seed, nmol=512)[:args.batch_mols]
    fbs = [m.flat() for m in mols]
    nmol = len(fbs)
    ctx = Context(0)
    st = torch.cuda.current_stream()
    sp = st.cuda_stream
    t = time.time()
    mplan = ctx.batch_plan(fbs, **MIXED)
    t_plan = time.time() - t
    n = mplan.n_unique
    off = mplan.mol_offsets()
    ms_stats = mplan.stats()
    log(f"[bench] headline: {nmol} molecules, {n:.3e} unique ERIs ({n * 8 / 1e9:.1f} GB), mixed plan {t_plan:.2f} s")
    # NaN first: a slot the graph leaves unwritten fails the error checks
    # below, which read the graph's output
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    mplan.capture(out.data_ptr())
    for _ in range(args.warmup):
        mplan.replay(sp)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(args.steps):
        mplan.replay(sp)
    e1.record(st)
    torch.cuda.synchronize()
    ms_mixed = e0.elapsed_time(e1) / args.steps
    value = n / (ms_mixed * 1e-3)
    log(f"[bench] batched mixed graph: {ms_mixed:.1f} ms/step, {value:.3e} unique ERIs/s")
    # errors of the mixed form (the graph's output): sampled slots against the
    # exact values, then every integral against the FP64 per-molecule plans
    err = {"unwritten_slots": sum(int(torch.isnan(out[i:i + (1 << 28)]).sum()) for i in range(0, n, 1 << 28)),
           "mixed_sampled": sampled_exact_errors(fbs, off, lambda idx: out[torch.from_numpy(idx).cuda()].cpu().numpy())}
    plans = [ctx.plan(fb) for fb in fbs]
    scratch = torch.empty(int(max(off[m + 1] - off[m] for m in range(nmol))), dtype=torch.float64, device="cuda")
    worst = torch.zeros((), dtype=torch.float64, device="cuda")
    for m, p in enumerate(plans):
        k = int(off[m + 1] - off[m])
        p.execute(scratch.data_ptr(), sp)
        worst = torch.maximum(worst, (out[int(off[m]):int(off[m]) + k] - scratch[:k]).abs().max())
    err["max_abs_err_mixed_vs_fp64_every_integral"] = float(worst)
    del scratch
    prof = mplan.profile(out.data_ptr(), sp)
    roof = roofline_from_profile(mplan, prof, ms_mixed, n)
    roof["profile_sum_ms"] = float(sum(prof))
    # plain path: per-molecule FP64 plans one after the other, same output
    base = out.data_ptr()

    def plain():
        for m, p in enumerate(plans):
            p.execute(base + 8 * int(off[m]), sp)
    ms_plain = timed(plain, 1, st)
    err["fp64_sampled"] = sampled_exact_errors(fbs, off, lambda idx: out[torch.from_numpy(idx).cuda()].cpu().numpy())
    for p in plans:
        p.close()
    mplan.close()
    bplan = ctx.batch_plan(fbs)
    bplan.capture(out.data_ptr())
    ms_fp64 = timed(lambda: bplan.replay(sp), 2, st)
    fl64 = plan_flops(bplan)
    b_stats = bplan.stats()
    bplan.close()
    del out
    torch.cuda.empty_cache()
    ctx.close()
    log(f"[bench] plain per-molecule {ms_plain:.1f} ms, batched fp64 {ms_fp64:.1f} ms, mixed {ms_mixed:.1f} ms: "
        f"{ms_plain / ms_mixed:.2f}x plain; mixed vs fp64 max {err['max_abs_err_mixed_vs_fp64_every_integral']:.2e}; "
        f"sampled vs exact: mixed {err['mixed_sampled']['max_abs_err_vs_exact']:.2e} "
        f"({err['mixed_sampled']['n_outside_parity_bar']} outside the bar), fp64 "
        f"{err['fp64_sampled']['max_abs_err_vs_exact']:.2e} ({err['fp64_sampled']['n_outside_parity_bar']})")
    unpruned = prim_quartets_unpruned(fbs)
    info = {"molecules": nmol, "unique_eris": int(n), "output_gb": n * 8 / 1e9, "plan_build_s": t_plan,
            "shell_quartets_computed_mixed": ms_stats["shell_quartets"],
            "primitive_quartets_computed_mixed": ms_stats["primitive_quartets"],
            "shell_quartets_fp64": b_stats["shell_quartets"],
            "primitive_quartets_fp64": b_stats["primitive_quartets"],
            "primitive_quartets_before_pair_prune": unpruned,
            "plain_per_molecule": {"ms": ms_plain, "value": n / (ms_plain * 1e-3)},
            "batched_fp64_graph": {"ms": ms_fp64, "value": n / (ms_fp64 * 1e-3), "algorithmic_flop": fl64,
                                   "achieved_tflops": fl64 / (ms_fp64 * 1e-3) / 1e12,
                                   "frac_fp64_peak": fl64 / (ms_fp64 * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS},
            "speedup_vs_plain": ms_plain / ms_mixed, "errors": err}
    return value, ms_mixed, roof, info, fbs


def leg_60bf(args):
    """The 60-function split-valence ethane-like molecule, plain FP64, chrr
    order: fused launch units, lanes balanced on measured unit times."""
    import torch
    from mi_amd import Context, configs, flops
    from mi_amd.hgp_tables import register_group
    from oracle import oracle as O  # noqa: F401  (checker of the fixture below)

    fb = configs.ethane_split_valence(args.seed).flat()
    ctx = Context(0)
    plan = ctx.plan(fb, order=1)
    n = plan.n_unique
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    for _ in range(20):
        plan.execute(out.data_ptr(), sptr)
    plan.balance(out.data_ptr(), sptr)
    ms = timed(lambda: plan.execute(out.data_ptr(), sptr), 200, stream)
    launches, units = plan.launches(), plan.units()
    acc = np.zeros(len(units) + 1)
    for _ in range(30):
        acc += np.array(plan.profile_units(out.data_ptr(), sptr))
    mean_ms = acc / 30
    sym = {}
    for i, (kind, nw, members) in enumerate(units):
        f = sum(flops.class_flops(launches[m][0], launches[m][1], launches[m][2]) for m in members)
        name = {1: f"hgp_rf{register_group(launches[members[0]][0])}_kernel", 2: "hgp_wf_kernel",
                3: "hgp_wgf_kernel"}.get(kind) or symbol_of(launches[members[0]])
        t, fl, nl = sym.get(name, (0.0, 0, 0))
        sym[name] = (t + mean_ms[i + 1], fl + f, nl + 1)
    name, (t, f, nl) = max(sym.items(), key=lambda kv: kv[1][0])
    g = np.load(os.path.join(ROOT, "tests", "golden", "ethane_split_valence_seed0.npz"))
    got = out.cpu().numpy()
    err = float(np.max(np.abs(got[g["slots"]] - g["values"]))) if args.seed == 0 else None
    plan.close()
    ctx.close()
    r = {"us_per_step": ms * 1e3, "value": n / (ms * 1e-3), "unique_eris": n, "launch_units": len(units) + 1,
         "dominant_symbol": name, "dominant_symbol_us": t * 1e3, "dominant_symbol_tflops": f / (t * 1e-3) / 1e12,
         "step_achieved_tflops": sum(x[1] for x in sym.values()) / (ms * 1e-3) / 1e12,
         "max_abs_err_vs_golden": err}
    if not args.no_cpu:
        r["cpu_baseline"] = cpu_baseline_single(fb, n, args.cpu_budget_s / 2)
    log(f"[bench] 60-function step: {ms * 1e3:.1f} us, {n / (ms * 1e-3):.3e} unique ERIs/s; dominant {name} "
        f"{t * 1e3:.1f} us")
    return r


# The reference's single-thread timings of its timing workloads (README.md:46-52,
# test/timing.jl:81-89: ethane in three named basis sets), seconds per call.
REFERENCE_TIMINGS_S = {"sto3g": {"readme_size": 16, "hgp": 0.041, "rys": 0.415},
                       "6-31G": {"readme_size": 30, "hgp": 0.202, "rys": 1.550},
                       "cc-pvdz": {"readme_size": 58, "hgp": 2.451, "rys": 25.94}}


def leg_reference_workloads():
    """Ethane in the reference's named basis sets through build_basis
    (test/timing.jl:81-89), HGP (all_twoe_ints_chrr) and Rys quadrature
    (all_twoe_ints_rys, iindex order) on the GPU: execute time with the basis
    resident (plan built once) and the one-shot call with host arrays in and
    out (mi_eri_unique, as the reference's call returns a host Vector), beside
    the reference's published single-thread times.  Errors: Rys against HGP on
    the GPU (different algorithms) and HGP against the FP64 oracle."""
    import torch
    from mi_amd import Context, build_basis, flatten
    from mi_amd.molecules import ethane
    from oracle import oracle as O

    ctx = Context(0)
    st = torch.cuda.current_stream()
    sp = st.cuda_stream
    res = {}
    for name, ref in REFERENCE_TIMINGS_S.items():
        fb = flatten(build_basis(ethane, name))
        r = {"nbf": fb.nbf, "unique_eris": fb.nbf * (fb.nbf + 1) * (fb.nbf ** 2 + fb.nbf + 2) // 8,
             "reference_single_thread_s": {"hgp": ref["hgp"], "rys": ref["rys"]}}
        outs = {}
        for meth, order, m in (("hgp", 1, 0), ("rys", 0, 1)):
            plan = ctx.plan(fb, order=order, method=m)
            out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
            ms = timed(lambda: plan.execute(out.data_ptr(), sp), 50, st)
            outs[meth] = out.cpu().numpy()
            plan.close()
            t = time.perf_counter()
            for _ in range(5):
                ctx.eri_unique(fb, order=order, method=m)
            one = (time.perf_counter() - t) / 5
            r[meth] = {"execute_ms": ms, "one_shot_ms": one * 1e3, "value": r["unique_eris"] / (ms * 1e-3),
                       "speedup_one_shot_vs_reference": ref[meth] / one}
        hgp_iindex = ctx.eri_unique(fb, order=0)
        r["max_abs_diff_rys_vs_hgp"] = float(np.abs(outs["rys"] - hgp_iindex).max())
        r["max_abs_err_hgp_vs_fp64_oracle"] = float(
            np.abs(outs["hgp"] - O.eri_unique_hgp(fb, order=1, nthreads=omp_threads())).max())
        res[name] = r
        log(f"[bench] ethane/{name} ({fb.nbf} functions): HGP {r['hgp']['execute_ms']:.3f} ms execute, "
            f"{r['hgp']['one_shot_ms']:.2f} ms one-shot (reference {ref['hgp']} s); Rys "
            f"{r['rys']['execute_ms']:.3f} ms execute, {r['rys']['one_shot_ms']:.2f} ms one-shot "
            f"(reference {ref['rys']} s); Rys vs HGP max {r['max_abs_diff_rys_vs_hgp']:.1e}")
    ctx.close()
    return res


def eri_spot_check(ctx, fb, nsets, seed=0):
    """The 920-function cluster's unique array (9e10 integrals) is never held:
    nsets random sets of 4 distinct shells, each a sub-basis with the
    cluster's own norms (its integrals are the cluster's), computed in one
    batched plan in plain FP64 and in the mixed, screened form.  Checker: the
    FP64 oracle, and where it is outside the bar of the exact value (HRR
    cancellation in the reference's shell order) the binary128 oracle.
    Screening and FP32 choices depend only on the quartet's pairs, so the
    mixed values are those the cluster's J/K build uses."""
    import torch
    from mi_amd import sub_basis
    from oracle import oracle as O

    rng = np.random.Generator(np.random.PCG64(seed + 77))
    subs = [sub_basis(fb, rng.choice(fb.nshell, 4, replace=False)) for _ in range(nsets)]
    ref64 = [O.eri_unique_hgp(sb, order=O.ORDER_CHRR, nthreads=omp_threads()) for sb in subs]
    sp = torch.cuda.current_stream().cuda_stream
    res = {"shell_sets": nsets, "shell_quartets": nsets * 55, "integrals": int(sum(r.size for r in ref64))}
    exact_cache = {}
    for name, kw in (("fp64", {}), ("mixed_screened", MIXED)):
        plan = ctx.batch_plan(subs, **kw)
        out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr(), sp)
        got = out.cpu().numpy()
        off = plan.mol_offsets()
        worst, nbad, n_arb, n_above = 0.0, 0, 0, 0
        for m, sb in enumerate(subs):
            g, r = got[off[m]:off[m + 1]], ref64[m]
            d = np.abs(g - r)
            bad = np.nonzero((d > 1e-12) & (d > 1e-10 * np.abs(r)))[0]
            dd = d.copy()
            if bad.size:
                if m not in exact_cache:
                    exact_cache[m] = {}
                need = [s for s in bad if s not in exact_cache[m]]
                if need:
                    ex = O.eri_slots_quad(sb, np.array(need), order=O.ORDER_CHRR, accurate=True,
                                          nthreads=omp_threads())
                    exact_cache[m].update(zip(need, ex))
                ex = np.array([exact_cache[m][s] for s in bad])
                de = np.abs(g[bad] - ex)
                ok = (de <= 1e-12) | (de <= 1e-10 * np.abs(ex))
                n_arb += int(ok.sum())
                nbad += int((~ok).sum())
                dd[bad] = de
            worst = max(worst, float(dd.max()))
            n_above += int(np.count_nonzero(dd > 1e-10))
        res[name] = {"max_abs_err_vs_exact": worst, "n_outside_parity_bar": nbad, "n_above_1e-10": n_above,
                     "n_where_fp64_oracle_is_off_and_gpu_is_not": n_arb}
        plan.close()
        del out
    res["max_abs_err_fp64_oracle_vs_exact"] = max(
        float(np.abs(np.array([ref64[m][s] for s in c]) - np.array(list(c.values()))).max())
        for m, c in exact_cache.items()) if exact_cache else 0.0
    res["max_abs_eri"] = max(float(np.abs(r).max()) for r in ref64)
    return res


def leg_jk(seed=0):
    """Direct J/K on the 40-atom cluster (920 functions, l <= 3): plain FP64
    (no screening) vs screening at 1e-12 vs mixed (FP32 VRR below 1e-8) with
    screening; errors of J and K against the plain FP64 GPU result, and the
    cluster's ERIs checked on 1820 sampled sets of 4 shells (100,100 shell
    quartets)."""
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
                           ("mixed_screened", MIXED, 2)):
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
    res["eri_spot_check"] = sc = eri_spot_check(ctx, fb, 1820, seed)
    log(f"[bench] J/K cluster ERI spot check: {sc['shell_quartets']} shell quartets, {sc['integrals']} integrals; "
        f"vs exact: fp64 max {sc['fp64']['max_abs_err_vs_exact']:.2e} ({sc['fp64']['n_outside_parity_bar']} outside "
        f"the bar), mixed max {sc['mixed_screened']['max_abs_err_vs_exact']:.2e} "
        f"({sc['mixed_screened']['n_above_1e-10']} above 1e-10); FP64 oracle vs exact "
        f"{sc['max_abs_err_fp64_oracle_vs_exact']:.2e}")
    ctx.close()
    res.update({"nbf": n, "unique_eris": nu, "screening": 1e-12,
                "speedup_mixed_screened_vs_plain": res["plain_fp64"]["ms"] / res["mixed_screened"]["ms"]})
    return res


def leg_serving(nreq, rate, nreq_plain, seed=0):
    """Poisson request schedule (2-60 atoms, split-valence or cluster recipe
    50/50; unique ERIs up to 120 functions, J/K beyond) replayed in real time
    through the continuous-batching server: the mixed form on all nreq
    requests, plain FP64 on the first nreq_plain (FP64 without screening takes
    seconds per large J/K request)."""
    from mi_amd import configs
    from mi_amd.serving import serve

    reqs = configs.serving_requests(seed=seed, nreq=nreq, rate_per_s=rate)
    out = {"requests": nreq, "rate_per_s": rate, "eri_requests": sum(r.kind == "eri" for r in reqs),
           "jk_requests": sum(r.kind == "jk" for r in reqs),
           "nbf_range": [min(r.workload.nbf for r in reqs), max(r.workload.nbf for r in reqs)]}
    res = {}
    for name, mixed, rs in (("mixed", True, reqs), ("plain_fp64", False, reqs[:nreq_plain])):
        r = serve(rs, mixed=mixed)
        res[name] = r
        lat = {k: np.array([r.latency_s[q.rid] for q in rs if q.kind == k]) * 1e3 for k in ("eri", "jk")}
        out[name] = {"requests": len(rs), "value": r.throughput, "unit": "unique ERIs/s",
                     "p50_ms": r.p(0.5) * 1e3, "p99_ms": r.p(0.99) * 1e3, "makespan_s": r.makespan_s,
                     "batches": r.stats["batches"], "peak_device_gb": r.stats["peak_bytes"] / 1e9,
                     "p50_ms_eri_requests": float(np.median(lat["eri"])) if lat["eri"].size else None,
                     "p99_ms_eri_requests": float(np.quantile(lat["eri"], 0.99)) if lat["eri"].size else None,
                     "p50_ms_jk_requests": float(np.median(lat["jk"])) if lat["jk"].size else None,
                     "p99_ms_jk_requests": float(np.quantile(lat["jk"], 0.99)) if lat["jk"].size else None,
                     **r.overlap()}
        log(f"[bench] serving {name} ({len(rs)} requests): p50 {r.p(0.5) * 1e3:.1f} ms p99 {r.p(0.99) * 1e3:.1f} ms "
            f"{r.throughput:.3e} unique ERIs/s, makespan {r.makespan_s:.2f} s, batches in flight mean "
            f"{r.overlap()['mean_batches_in_flight']:.2f} max {r.overlap()['max_batches_in_flight']}")
    err = {"eri": 0.0, "jk": 0.0}
    for q in reqs[:nreq_plain]:
        a = [t.cpu().numpy() for t in res["plain_fp64"].outputs[q.rid]]
        b = [t.cpu().numpy() for t in res["mixed"].outputs[q.rid]]
        k = 0 if q.kind == "eri" else 1
        err[q.kind] = max(err[q.kind], max(float(np.abs(x - y).max()) for x, y in zip(a[k:], b[k:])))
    out["mixed"]["max_abs_err_eri_vs_fp64"] = err["eri"]
    out["mixed"]["max_abs_err_jk_vs_fp64"] = err["jk"]
    del res
    return out


def leg_served_batch(fbs, plain_ms):
    """The headline's 512 molecules submitted at once as ERI requests (mixed
    form) to the continuous-batching server: throughput of the served form,
    to set beside the plain per-molecule path and the batched graph."""
    from mi_amd.configs import Request
    from mi_amd.serving import serve

    class _W:  # a workload whose flattened basis is already built
        def __init__(self, fb):
            self._fb = fb
            self.nbf = fb.nbf

        def flat(self):
            return self._fb

    reqs = [Request(i, 0.0, _W(fb), "eri") for i, fb in enumerate(fbs)]
    r = serve(reqs, mixed=True, max_inflight=8, max_batch=64, hbm_budget=40 << 30)
    n = r.unique_eris
    out = {"requests": len(reqs), "value": r.throughput, "makespan_s": r.makespan_s, "batches": r.stats["batches"],
           "p50_ms": r.p(0.5) * 1e3, "p99_ms": r.p(0.99) * 1e3, "peak_device_gb": r.stats["peak_bytes"] / 1e9,
           "speedup_vs_plain": (plain_ms * 1e-3) / r.makespan_s, **r.overlap()}
    log(f"[bench] served batch: {len(reqs)} requests, {n:.3e} ERIs in {r.makespan_s:.2f} s = {r.throughput:.3e}/s "
        f"({out['speedup_vs_plain']:.2f}x the plain path), batches in flight mean {out['mean_batches_in_flight']:.2f}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget-s", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="only the headline workload")
    ap.add_argument("--batch-mols", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=128)
    ap.add_argument("--serve-plain-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--profile-headline", action="store_true",
                    help="only execute the headline plan once after a warm-up, one stream (for rocprofv3), and exit")
    ap.add_argument("--profile-batch", type=int, default=0, metavar="NMOL",
                    help="only execute one FP64 batched plan of NMOL molecules (seed 1), one stream, and exit")
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
    if args.profile_headline or args.profile_batch:
        import torch
        from mi_amd import Context, configs
        ctx = Context(0)
        ctx.set_streams(1)
        if args.profile_headline:
            fbs = [m.flat() for m in configs.batch_molecules(args.seed, nmol=512)[:args.batch_mols]]
            plan = ctx.batch_plan(fbs, **MIXED)
        else:
            fbs = [m.flat() for m in configs.batch_molecules(1, nmol=args.profile_batch)]
            plan = ctx.batch_plan(fbs)
        out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
        for _ in range(2 if args.profile_headline else 1):
            plan.execute(out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        print(json.dumps({"molecules": len(fbs), "unique_eris": plan.n_unique, "executes": 2 if args.profile_headline else 1,
                          "launches": [["".join("spdf"[x] for x in l[0]), l[4], l[1], l[2]]
                                       for l in plan.launches()]}), flush=True)
        plan.close()
        ctx.close()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)

    value, ms_step, roof, info, fbs = headline(args)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "unique ERIs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 VRR for quartets with V_ab V_cd < 1e-8)",
        "data": "synthetic (seeded bases and geometries)",
        "config": {"workload": f"{info['molecules']} molecules of 3-24 atoms (split-valence + d, seed {args.seed}), "
                               "one batched plan replayed as a captured graph, mixed form (Schwarz cut 1e-12, "
                               "FP32 VRR below tau = 1e-8), every unique ERI of every molecule in chrr order",
                   "molecules": info["molecules"], "unique_eris": info["unique_eris"], "output_gb": info["output_gb"],
                   "precision": "mixed", "order": "chrr", "schwarz_cut": 1e-12, "tau": 1e-8},
        "speedup_vs_plain": info["speedup_vs_plain"],
        "plain": info["plain_per_molecule"],
        "batched_fp64": info["batched_fp64_graph"],
        "errors": info["errors"],
        "work": {k: info[k] for k in ("shell_quartets_computed_mixed", "primitive_quartets_computed_mixed",
                                      "shell_quartets_fp64", "primitive_quartets_fp64",
                                      "primitive_quartets_before_pair_prune", "plan_build_s")},
        "roofline": roof,
    }
    if not args.no_cpu:
        log("[bench] timing CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline_batch(fbs, args.cpu_budget_s)
    if not args.no_legs:
        import torch
        legs = {}
        for name, fn in (("ethane_split_valence_60bf", lambda: leg_60bf(args)),
                         ("reference_timing_workloads", leg_reference_workloads),
                         ("jk_cluster_40atoms", leg_jk),
                         ("serving", lambda: leg_serving(args.serve_requests, args.serve_rate,
                                                         args.serve_plain_requests)),
                         ("served_batch", lambda: leg_served_batch(fbs, info["plain_per_molecule"]["ms"]))):
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
