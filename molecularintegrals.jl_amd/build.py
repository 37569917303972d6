"""Build libmi_hgp.so in-tree for gfx950.

Steps: run csrc/gen_hgp.py (generated per-class kernels and op tables into
csrc/gen/) and csrc/gen_rys.c (Rys quadrature tables into csrc/gen/), then
hipcc the HIP kernels and the host/C-ABI sources into one shared library next
to this file.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmi_hgp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-Wno-unused-value", "-munsafe-fp-atomics"]


def run(cmd):
    t = time.time()
    subprocess.check_call(cmd)
    return time.time() - t


def build(verbose: bool = True) -> str:
    gen_dir = os.path.join(CSRC, "gen")
    run([sys.executable, os.path.join(CSRC, "gen_hgp.py"), "--out", gen_dir])
    # Rys quadrature tables: gen_rys.c (binary128 Gauss rules, gcc + libquadmath)
    gen_rys = os.path.join(gen_dir, "gen_rys")
    run(["gcc", "-O2", "-o", gen_rys, os.path.join(CSRC, "gen_rys.c"), "-lquadmath", "-lm"])
    run([gen_rys, "--out", os.path.join(gen_dir, "rys_table.inc")])
    objs = []
    jobs = []
    # the fused register kernels: mi_fused.hip once per fusion group (the
    # heaviest groups first, they take longest to compile)
    units = [("mi_fused.hip", [f"-DMI_FUSED_GROUP={g}"], f"mi_fused_{g}") for g in (4, 3, 2, 1, 0)]
    units += [(src, [], src) for src in ("mi_kernels.hip", "mi_rys.hip", "mi_host.cpp", "mi_server.cpp")]
    for src, defs, stem in units:
        obj = os.path.join(gen_dir, stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [HIPCC] + COMMON + defs + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append(subprocess.Popen(cmd))
        objs.append(obj)
    for j in jobs:
        if j.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
