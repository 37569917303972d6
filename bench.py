#!/usr/bin/env python3
"""Throughput benchmark of the HGP two-electron integral path on one MI355X.

A step is one full pass of the hot path over the 60-function split-valence
ethane-like workload (seed 0): the shell-pair kernel plus every per-class ERI
kernel, writing all 1,675,365 unique (ij|kl) in the chrr order into a
device-resident buffer.  Inputs (basis topology and primitives) are resident
in HBM before the timed region.

Prints ONE JSON line.  Extra keys:
  roofline      dominant kernel: algorithmic FP64 FLOP per launch (FLOP model
                of mi_amd/flops.py) / its mean HIP-event duration, against the
                FP64 vector peak (the recurrences do not use MFMA).
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
MEASURED_TRAFFIC_BYTES = {"hgp_w_kernel": 1230314}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
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
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)

    import torch
    from mi_amd import Context, configs, flops

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

    # Per-launch durations (HIP events between launches, same stream).  The
    # roofline is taken for the dominant kernel SYMBOL: hgp_r_<class> is one
    # class, hgp_w_kernel / hgp_wg_kernel serve several classes, so their time
    # and FLOPs are summed over the launches of one step and the mean launch
    # duration is total / launches (what rocprofv3 --stats reports per symbol).
    launches = plan.launches()
    nprof = max(5, min(args.steps, 50))
    acc = np.zeros(len(launches) + 1)
    for _ in range(nprof):
        acc += np.array(plan.profile(out.data_ptr(), sptr))
    mean_ms = acc / nprof
    per = []
    step_flops = 0
    sym = {}
    for i, (cls, nqs, nqp, nblk, kern) in enumerate(launches):
        f = flops.class_flops(cls, nqs, nqp)
        step_flops += f
        per.append((mean_ms[i + 1], cls, f, nqs, nqp, kern))
        name = {"R": "hgp_r_" + "".join("spdf"[x] for x in cls), "W": "hgp_w_kernel", "G": "hgp_wg_kernel"}[kern[0]]
        t, fl, nl = sym.get(name, (0.0, 0, 0))
        sym[name] = (t + mean_ms[i + 1], fl + f, nl + 1)
    name, (dom_ms_total, dom_flops, dom_n) = max(sym.items(), key=lambda kv: kv[1][0])
    achieved = dom_flops / (dom_ms_total * 1e-3) / 1e12
    for ms_, cls, f, nqs, nqp, kern in sorted(per, key=lambda x: -x[0])[:8]:
        log(f"[bench]   {''.join('spdf'[x] for x in cls)} {kern} {ms_ * 1e3:8.2f} us "
            f"{f / (ms_ * 1e-3) / 1e12:7.3f} TFLOP/s  quartets={nqs} prim={nqp}")
    for nm, (t, fl, nl) in sorted(sym.items(), key=lambda kv: -kv[1][0])[:5]:
        log(f"[bench]   symbol {nm}: {nl} launch(es), {t * 1e3:.2f} us/step, {fl / (t * 1e-3) / 1e12:.3f} TFLOP/s")
    log(f"[bench]   pair kernel {mean_ms[0] * 1e3:.2f} us; sum of launches {mean_ms.sum() * 1e3:.1f} us")

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
                     "step_output_write_gbs": n * 8 / (ms_per_step * 1e-3) / 1e9},
        "max_abs_err_vs_golden": err,
    }
    if not args.no_cpu:
        log("[bench] timing CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(fb, n, args.cpu_budget_s)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
