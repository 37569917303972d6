#!/usr/bin/env python3
"""Generator for the HGP ERI device code (gfx950).

It symbolically runs the reference's recurrences and emits HIP source:

* VRR -- HGP eq. 6 exactly as ``vrr!`` applies it (rpmuller/MolecularIntegrals.jl
  src/HGP.jl:135-242): the raised index is lowered along
  ``shift_direction`` (first argmax of the Cartesian powers,
  src/Basis.jl:346-354), a-build first with c = 0, then the c-build.
* HRR -- ``chrr`` (src/HGP.jl:83-116): (a,b+1_j| = (a+1_j,b| + AB_j (a,b|,
  then (ab|c,d+1_j) = (ab|c+1_j,d) + CD_j (ab|c,d), j = shift_direction.

Only the entries the target block needs are generated (the reference fills the
whole cumulative tables).  Two outputs:

``hgp_rkernels.inc``  one straight-line kernel per "register" class: a lane
    group of G lanes owns a shell quartet, splits its primitive quartets,
    keeps VRR intermediates and contracted (e0|f0) in VGPRs, reduces across the
    group with cross-lane shuffles, runs HRR in registers and scatters.
``hgp_wtables.inc``   op tables for every class, consumed by the generic
    wave-cooperative kernel (one wavefront per shell quartet, VRR/HRR levels
    staged in LDS, lanes parallel within a level).

Cartesian component order: src/Basis.jl:269-279 for L<=4 (f: xyy, xzz, xyz),
descending I then J for L>=5 (HRR intermediates only).
"""
from __future__ import annotations

import argparse
import functools
import os
import sys
from collections import OrderedDict


sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from mi_amd.hgp_tables import (AO2M, LMAX_SHELL, HrrGraph, VrrGraph, cname, lval, nao, ncart,  # noqa: E402
                               quartet_classes, sdir, shell_aos, sidx, sidx_plus, NRGROUP, register_group,
                               is_swapped_class)

# ---------------------------------------------------------------- R emitter

XYZ = "xyz"


def vname(e, f, m):
    return f"v{e}_{f}_{m}"


def vrr_block(g, ctype, indent):
    """Straight-line VRR for one primitive quartet in arithmetic type ctype;
    returns (coefficient definitions, statements, target accumulations)."""
    uses = set()
    body = []
    cast = "" if ctype == "double" else "(float)"
    for (e, f) in g.nodes:
        rec = g.recurrence(e, f)
        for m in range(g.maxm[(e, f)] + 1):
            if rec[0] == "F":
                body.append(f"{indent}const {ctype} {vname(e, f, m)} = {cast}(pref * F[{m}]);")
            elif rec[0] == "A":
                _, i, e1, e2, n1 = rec
                x = XYZ[i]
                uses.update({f"PA{x}", f"WP{x}"})
                expr = f"PA{x} * {vname(e1, 0, m)} + WP{x} * {vname(e1, 0, m + 1)}"
                if e2 >= 0:
                    uses.update({"oo2z", "eta_s"})
                    coef = "oo2z" if n1 == 1 else f"({cast}{float(n1)} * oo2z)"
                    expr += f" + {coef} * ({vname(e2, 0, m)} - eta_s * {vname(e2, 0, m + 1)})"
                body.append(f"{indent}const {ctype} {vname(e, f, m)} = {expr};")
            else:
                _, i, f1, f2, e1, ne, nf = rec
                x = XYZ[i]
                uses.update({f"QC{x}", f"WQ{x}"})
                expr = f"QC{x} * {vname(e, f1, m)} + WQ{x} * {vname(e, f1, m + 1)}"
                if e1 >= 0:
                    uses.add("oo2s")
                    coef = "oo2s" if ne == 1 else f"({cast}{float(ne)} * oo2s)"
                    expr += f" + {coef} * {vname(e1, f1, m + 1)}"
                if f2 >= 0:
                    uses.update({"oo2e", "zeta_s"})
                    coef = "oo2e" if nf == 1 else f"({cast}{float(nf)} * oo2e)"
                    expr += f" + {coef} * ({vname(e, f2, m)} - zeta_s * {vname(e, f2, m + 1)})"
                body.append(f"{indent}const {ctype} {vname(e, f, m)} = {expr};")
    defs = []
    for x in XYZ:
        if f"PA{x}" in uses:
            defs.append(f"{indent}const {ctype} PA{x} = {cast}bp.PA{x};")
        if f"WP{x}" in uses:
            defs.append(f"{indent}const {ctype} WP{x} = {cast}(-eta * inv_s * PQ{x});")
        if f"QC{x}" in uses:
            defs.append(f"{indent}const {ctype} QC{x} = {cast}kp.PA{x};")
        if f"WQ{x}" in uses:
            defs.append(f"{indent}const {ctype} WQ{x} = {cast}(zeta * inv_s * PQ{x});")
    if "oo2z" in uses:
        defs.append(f"{indent}const {ctype} oo2z = {cast}bp.oo2z;")
    if "oo2e" in uses:
        defs.append(f"{indent}const {ctype} oo2e = {cast}kp.oo2z;")
    if "oo2s" in uses:
        defs.append(f"{indent}const {ctype} oo2s = {cast}(0.5 * inv_s);")
    if "eta_s" in uses:
        defs.append(f"{indent}const {ctype} eta_s = {cast}(eta * inv_s);")
    if "zeta_s" in uses:
        defs.append(f"{indent}const {ctype} zeta_s = {cast}(zeta * inv_s);")
    accs = [f"{indent}acc{e}_{f} += (double){vname(e, f, 0)};" for (e, f) in g.targets]
    return defs, body, accs


# Boys grid of a register class: staged into LDS per block (30.8 KB) for the
# light classes (top order <= 2), whose primitive loops are short and read it
# most; read through L1/L2 for the others, whose many VGPRs already limit the
# blocks per CU and whose blocks would otherwise hold the LDS while their
# slowest wavefront finishes.
BOYS_LDS_MAX_MMAX = 3


def boys_in_lds(c):
    return VrrGraph(*c).mmax <= BOYS_LDS_MAX_MMAX


def emit_rkernel(c):
    la, lb, lc, ld = c
    g = VrrGraph(*c)
    h = HrrGraph(*c)
    amax, cmax, mmax = g.amax, g.cmax, g.mmax
    name = cname(c)
    L = []
    w = L.append
    w(f"// ---- class ({name[0]}{name[1]}|{name[2]}{name[3]}): amax={amax} cmax={cmax} mmax={mmax}, "
      f"{g.count_entries()} VRR entries, {len(g.targets)} contracted (e0|f0), {len(h.nodes)} HRR nodes")
    # The class body takes its block index within the class launch (bid) and
    # the block's Boys slice area, so the single-class kernel hgp_r_<class> and
    # the fused multi-class kernel hgp_rf_kernel run the same code.
    w(f"__device__ __forceinline__ void hgp_r_body_{name}(const EriArgs& args, double* sboys, int bid) {{")
    if boys_in_lds(c):
        w(f"  stage_boys_slice(sboys, args.boys, {mmax - 1});")
        w(f"  const double* btab = sboys;")
    else:
        w(f"  const double* btab = args.boys + (size_t){mmax - 1} * BOYS_NROW * 8;  // read through L1/L2")
    w(f"  QuartetCtx qc;")
    # Lanes without a quartet (past the task's end, or screened) return, except
    # in the J/K build: there every lane joins the wavefront's run reduction
    # (jk_block_regs) with a zero block.
    w(f"  bool live = decode_quartet(args, qc, bid);")
    w(f"  if (!live && args.mode != MI_MODE_JK) return;")
    w(f"  const PairTopo P = args.topo[qc.bra];")
    w(f"  const PairTopo Q = args.topo[qc.ket];")
    w(f"  const int qmode = live ? quartet_mode(args, qc) : 2;  // 0 FP64 VRR, 1 FP32 VRR, 2 screened")
    w("  if (qmode == 2) {  // screened: output pre-zeroed")
    w("    if (args.mode != MI_MODE_JK) return;")
    w("    live = false;")
    w("  }")
    for (e, f) in g.targets:
        w(f"  double acc{e}_{f} = 0.0;")
    w(f"  const int nu = P.npp, nv = Q.npp, ntot = live ? nu * nv : 0;")
    w(f"  const double Ax = args.shell_xyz[3 * P.shA], Ay = args.shell_xyz[3 * P.shA + 1], "
      f"Az = args.shell_xyz[3 * P.shA + 2];")
    w(f"  const double Cx = args.shell_xyz[3 * Q.shA], Cy = args.shell_xyz[3 * Q.shA + 1], "
      f"Cz = args.shell_xyz[3 * Q.shA + 2];")
    w(f"  const int G = qc.G;")
    if mmax > 1:
        # FP32 VRR only where every live quartet of the wavefront qualifies:
        # the two VRR paths of a mixed wavefront would run one after the other
        # (non-packed FP32 issues at the FP64 rate, so FP32 saves nothing there)
        w(f"  const bool use32 = __all(qmode == 1 || !live);")
    w(f"  const double rnv = rcp_int(Q.npp);")
    w(f"  int v, dv;")
    w(f"  int u = udiv_rcp(qc.g, Q.npp, rnv, v);")
    w(f"  const int du = udiv_rcp(G, Q.npp, rnv, dv);")
    w(f"  for (int t = qc.g; t < ntot; t += G) {{")
    w(f"    const PrimPair bp = args.pp[P.ppoff + u];")
    w(f"    const PrimPair kp = args.pp[Q.ppoff + v];")
    w(f"    v += dv; u += du; if (v >= nv) {{ v -= nv; ++u; }}")
    w(f"    const double zeta = bp.zeta, eta = kp.zeta;")
    w(f"    const double rs = rsqrt_nr(zeta + eta), inv_s = rs * rs;")
    w(f"    const double PQx = (Ax + bp.PAx) - (Cx + kp.PAx), PQy = (Ay + bp.PAy) - (Cy + kp.PAy),")
    w(f"                 PQz = (Az + bp.PAz) - (Cz + kp.PAz);")
    w(f"    const double T = zeta * eta * inv_s * (PQx * PQx + PQy * PQy + PQz * PQz);")
    w(f"    const double pref = bp.K * kp.K * rs;")
    w(f"    double F[{mmax}];")
    w(f"    boys_eval<{mmax}>(T, btab, F);")
    if mmax > 1:
        w(f"    if (use32) {{  // FP32 VRR (Boys and prefactor stay FP64)")
        d, bdy, acc = vrr_block(g, "float", "      ")
        L.extend(d + bdy + acc)
        w(f"    }} else {{")
        d, bdy, acc = vrr_block(g, "double", "      ")
        L.extend(d + bdy + acc)
        w(f"    }}")
    else:
        d, bdy, acc = vrr_block(g, "double", "    ")
        L.extend(d + bdy + acc)
    w("  }")
    # cross-lane reduction within the lane group (fixed butterfly: deterministic)
    w("  for (int s = G >> 1; s > 0; s >>= 1) {")
    for (e, f) in g.targets:
        w(f"    acc{e}_{f} += __shfl_xor(acc{e}_{f}, s, 64);")
    w("  }")
    diag = [(e, f) for (e, f) in g.targets if e == f]
    if diag:
        w("  if (args.mode == MI_MODE_SCHWARZ && qc.g == 0) {  // bound of the VRR outputs (e0|e0)")
        w("    double vmax = 0.0;")
        for (e, f) in diag:
            w(f"    vmax = fmax(vmax, fabs(acc{e}_{f}));")
        w("    record_vmax(&args.geom[qc.bra], vmax);")
        w("  }")
    # HRR (always FP64)
    if lb or ld:
        w("  const PairGeom PG = args.geom[qc.bra];")
        w("  const PairGeom QG = args.geom[qc.ket];")
    for n in h.nodes:
        kind, s1, s2, j = h.deps(n)
        a, b, cc, d = n
        nm = f"h{a}_{b}_{cc}_{d}"
        if kind == "acc":
            w(f"  const double {nm} = acc{a}_{cc};")
        elif kind == "bra":
            w(f"  const double {nm} = h{s1[0]}_{s1[1]}_{s1[2]}_{s1[3]} + PG.AB{XYZ[j]} * h{s2[0]}_{s2[1]}_{s2[2]}_{s2[3]};")
        else:
            w(f"  const double {nm} = h{s1[0]}_{s1[1]}_{s1[2]}_{s1[3]} + QG.AB{XYZ[j]} * h{s2[0]}_{s2[1]}_{s2[2]}_{s2[3]};")
    w(f"  double sA[{ncart(la)}], sB[{ncart(lb)}], sC[{ncart(lc)}], sD[{ncart(ld)}];")
    w(f"  load_scales<{ncart(la)}, {ncart(lb)}, {ncart(lc)}, {ncart(ld)}>(args, P, Q, sA, sB, sC, sD);")
    w("  const ScatterCtx sc = make_scatter(args, qc, P, Q);")
    # scaled block values, element e = ((ma nb + mb) nc + mc) nd + md
    nout = 0
    for ma, a in enumerate(shell_aos(la)):
        for mb, b in enumerate(shell_aos(lb)):
            for mc, cc in enumerate(shell_aos(lc)):
                for md, d in enumerate(shell_aos(ld)):
                    w(f"  const double o{nout} = h{a}_{b}_{cc}_{d} * (sA[{ma}] * sB[{mb}]) * (sC[{mc}] * sD[{md}]);")
                    nout += 1
    # Coulomb/exchange: contract the scaled block in registers, a few atomics per quartet
    w("  if (args.mode == MI_MODE_JK) {")
    w(f"    const double ov[{nout}] = {{")
    for k in range(0, nout, 8):
        w("      " + ", ".join(f"o{e}" for e in range(k, min(nout, k + 8))) + ",")
    w("    };")
    w(f"    jk_block_regs<{ncart(la)}, {ncart(lb)}, {ncart(lc)}, {ncart(ld)}>(args, sc, live, ov);")
    w("    return;")
    w("  }")
    # Scatter: the G lanes of the quartet hold the same block; lane g of W =
    # min(G, 8) takes elements e = e0 + g (value picked by lane, no divergent
    # per-element code), so a wavefront issues nout / W slot computations.
    tmpl = f"<{ncart(lb)}, {ncart(lc)}, {ncart(ld)}>"
    w("  const int gl = qc.g;")
    w("  switch (qc.G) {  // uniform per wavefront (a wavefront holds one task)")
    for W in (1, 2, 4, 8):
        w(f"    {'default' if W == 8 else 'case ' + str(W)}: {{")
        if W > 1:
            w(f"      if (gl >= {W}) break;")
        for e0 in range(0, nout, W):
            k = min(W, nout - e0)
            if W == 1:
                w(f"      scatter_block_elem{tmpl}(args, sc, {e0}, o{e0});")
                continue
            sel = f"o{e0 + k - 1}"
            for j in range(k - 2, -1, -1):
                sel = f"gl == {j} ? o{e0 + j} : {sel}"
            if k < W:
                w(f"      if (gl < {k}) scatter_block_elem{tmpl}(args, sc, {e0} + gl, {sel});")
            else:
                w(f"      scatter_block_elem{tmpl}(args, sc, {e0} + gl, {sel});")
        w("      break;")
        w("    }")
    w("  }")
    w("}")
    w("#ifndef MI_R_BODIES_ONLY")
    w(f"__global__ void __launch_bounds__(256) hgp_r_{name}(const EriArgs args) {{")
    if boys_in_lds(c):
        w(f"  __shared__ double sboys[BOYS_NROW * 8];")
        w(f"  hgp_r_body_{name}(args, sboys, blockIdx.x);")
    else:
        w(f"  hgp_r_body_{name}(args, nullptr, blockIdx.x);")
    w("}")
    w("#endif")
    return "\n".join(L), g, h


# ---------------------------------------------------------------- W tables


def w_tables(c):
    """Op tables for the wave-cooperative kernel.

    VRR: the work area holds [e|f]^(m) for the pruned (e,f) nodes, m contiguous,
    nodes in level order; items enumerate the same order, so item k of the class
    writes slot vt0 + k (targets are implicit).  Items are returned unpacked as
    (s1, s2, s3, dir, n1, n2, kind, has2, has3); pack16/pack32 encode them.
    kind 0 (a-build): PA_i [s1] + WP_i [s1+1] + n1/(2 zeta) ([s2] - eta/s [s2+1])
    kind 1 (c-build): QC_i [s1] + WQ_i [s1+1] + n1/(2 s) [s3+1]
                      + n2/(2 eta) ([s2] - zeta/s [s2+1])
    HRR: slots 0..nacc-1 are the contracted (e0|f0) (accsrc[x] = VRR slot
    of its m=0 value, bit 31 set for a diagonal e == f node); item k writes slot
    nacc + k as (s1, s2, dir, ket)  ->  [s1] + X_dir [s2].
    Out: (hrr slot, ma, mb, mc, md).
    """
    la, lb, lc, ld = c
    g = VrrGraph(*c)
    h = HrrGraph(*c)
    base = {}
    off = 0
    for ef in g.nodes:
        base[ef] = off
        off += g.maxm[ef] + 1
    nvrr = off
    assert g.nodes[0] == (0, 0)
    vt0 = g.maxm[(0, 0)] + 1
    levels = sorted(set(g.level(*ef) for ef in g.nodes if ef != (0, 0)))
    vitems, vlev = [], []
    for lev in levels:
        vlev.append(len(vitems))
        for ef in g.nodes:
            if ef == (0, 0) or g.level(*ef) != lev:
                continue
            e, f = ef
            rec = g.recurrence(e, f)
            for m in range(g.maxm[ef] + 1):
                assert base[ef] + m == vt0 + len(vitems)
                if rec[0] == "A":
                    _, i, e1, e2, n1 = rec
                    s1 = base[(e1, 0)] + m
                    has2 = e2 >= 0
                    s2 = base[(e2, 0)] + m if has2 else 0
                    vitems.append((s1, s2, 0, i, n1, 0, 0, int(has2), 0))
                else:
                    _, i, f1, f2, e1, ne, nf = rec
                    s1 = base[(e, f1)] + m
                    has3 = e1 >= 0
                    s3 = base[(e1, f1)] + m if has3 else 0
                    has2 = f2 >= 0
                    s2 = base[(e, f2)] + m if has2 else 0
                    vitems.append((s1, s2, s3, i, ne, nf, 1, int(has2), int(has3)))
    vlev.append(len(vitems))
    acc_nodes = [n for n in h.nodes if h.deps(n)[0] == "acc"]
    hlevels = sorted(set(h.level(n) for n in h.nodes if h.level(n) > 0))
    hbase = {n: k for k, n in enumerate(acc_nodes)}
    order = []
    for lev in hlevels:
        for n in h.nodes:
            if h.level(n) == lev:
                hbase[n] = len(acc_nodes) + len(order)
                order.append(n)
    nhrr = len(acc_nodes) + len(order)
    accsrc = [base[(n[0], n[2])] | ((1 << 31) if n[0] == n[2] else 0) for n in acc_nodes]
    # Rys kernel: Cartesian powers of e and f of each contracted (e0|f0), 3 bits each
    rysexp = []
    for n in acc_nodes:
        (ex, ey, ez), (fx, fy, fz) = AO2M[n[0]], AO2M[n[2]]
        rysexp.append(ex | ey << 3 | ez << 6 | fx << 9 | fy << 12 | fz << 15)
    hitems, hlev = [], []
    for lev in hlevels:
        hlev.append(len(hitems))
        for n in order:
            if h.level(n) != lev:
                continue
            kind, s1, s2, j = h.deps(n)
            assert hbase[n] == len(acc_nodes) + len(hitems)
            hitems.append((hbase[s1], hbase[s2], j, 1 if kind == "ket" else 0))
    hlev.append(len(hitems))
    outs = []
    for ma, a in enumerate(shell_aos(la)):
        for mb, b in enumerate(shell_aos(lb)):
            for mc, cc in enumerate(shell_aos(lc)):
                for md, d in enumerate(shell_aos(ld)):
                    outs.append((hbase[(a, b, cc, d)], ma, mb, mc, md))
    assert len(outs) <= nvrr  # the Coulomb/exchange path stages the block in the VRR area
    desc = dict(la=la, lb=lb, lc=lc, ld=ld, mmax=g.mmax, nvrr=nvrr, nhrr=nhrr, vt0=vt0,
                vlev=vlev, hlev=hlev, nacc=len(accsrc), nout=len(outs))
    return desc, vitems, accsrc, hitems, outs, rysexp


# Per-primitive coefficient block of the LDS wave kernel (doubles, W_CF of
# them per wavefront): the VRR items index it, so one formula serves both
# builds without selects.
CF_PA, CF_WP, CF_QC, CF_WQ = 0, 3, 6, 9      # + direction
CF_ETA_S, CF_ZETA_S = 12, 13                 # eta/(zeta+eta), zeta/(zeta+eta)
CF_N_OO2Z, CF_N_OO2E, CF_N_OO2S = 16, 24, 32  # + n: n/(2 zeta), n/(2 eta), n/(2(zeta+eta)), n = 0..7


def vrr_item_w(s1, s2, s3, dr, n1, n2, kind, h2, h3):
    """LDS-kernel VRR item (4 words): value = cf[A] v[s1] + cf[B] v[s1+1]
    + cf[C] v[s3+1] + cf[D] (v[s2] - cf[E] v[s2+1]), slots as byte offsets.
      w0 = 8 s1 | A << 20 | B << 26,  w1 = 8 s2 | C << 20 | D << 26,  w2 = 8 s3 | E << 20
    a-build: A = PA_i, B = WP_i, C = 0, D = n1/(2 zeta), E = eta/s
    c-build: A = QC_i, B = WQ_i, C = n1/(2 s), D = n2/(2 eta), E = zeta/s"""
    assert max(s1, s2, s3) < (1 << 17) and n1 < 8 and n2 < 8
    if kind == 0:
        A, B, C, D, E = CF_PA + dr, CF_WP + dr, CF_N_OO2S, CF_N_OO2Z + (n1 if h2 else 0), CF_ETA_S
    else:
        A, B = CF_QC + dr, CF_WQ + dr
        C, D, E = CF_N_OO2S + (n1 if h3 else 0), CF_N_OO2E + (n2 if h2 else 0), CF_ZETA_S
    return [8 * s1 | A << 20 | B << 26, 8 * s2 | C << 20 | D << 26, 8 * s3 | E << 20, 0]


def pack16(vi, hi, ou, am):
    """LDS-kernel blob: VRR items (vrr_item_w, 4 words), then uint2 HRR items
    with 16-bit slot fields, out words and acc sources:
      HRR  x = s1 | s2 << 16, y = dir | ket << 2
      out  slot | ma << 16 | mb << 20 | mc << 24 | md << 28;  acc as is."""
    words = []
    for it in vi:
        words += vrr_item_w(*it)
    for s1, s2, dr, ket in hi:
        words += [s1 | (s2 << 16), dr | (ket << 2)]
    words += [o | (ma << 16) | (mb << 20) | (mc << 24) | (md << 28) for o, ma, mb, mc, md in ou]
    words += list(am)
    while len(words) % 4:
        words.append(0)
    return words


def pack32(vi, hi, ou, am, rx):
    """Global-kernel tables, 32-bit slot fields; returns (words, offsets).
      VRR  uint4 (s1, s2, s3, dir | n1 << 2 | n2 << 5 | kind << 8 | has2 << 9 | has3 << 10)
      HRR  uint4 (s1, s2, dir | ket << 2, 0)
      out  uint2 (slot, ma | mb << 4 | mc << 8 | md << 12)
      acc  uint2 (VRR slot, diagonal flag)
      rys  uint32 per contracted (e0|f0): ex | ey << 3 | ez << 6 | fx << 9 | fy << 12 | fz << 15"""
    w = []
    gv = 0
    for s1, s2, s3, dr, n1, n2, kind, h2, h3 in vi:
        w += [s1, s2, s3, dr | (n1 << 2) | (n2 << 5) | (kind << 8) | (h2 << 9) | (h3 << 10)]
    gh = len(w)
    for s1, s2, dr, ket in hi:
        w += [s1, s2, dr | (ket << 2), 0]
    go = len(w)
    for o, ma, mb, mc, md in ou:
        w += [o, ma | (mb << 4) | (mc << 8) | (md << 12)]
    ga = len(w)
    for a_ in am:
        w += [a_ & 0x7FFFFFFF, a_ >> 31]
    gr = len(w)
    w += list(rx)
    while len(w) % 4:
        w.append(0)
    return w, (gv, gh, go, ga, gr)


# ---------------------------------------------------------------- driver

# Classes compiled as straight-line register kernels.  The rest use the
# table-driven wave-cooperative kernel.  Chosen by contracted-accumulator count
# and VRR size so the register kernels stay within the VGPR budget.
def is_register_class(c):
    if is_swapped_class(c):  # rare (see pair_classes): table-driven kernels only
        return False
    g = VrrGraph(*c)
    return len(g.targets) <= 100 and g.count_entries() <= 320


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--lmax", type=int, default=LMAX_SHELL)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    classes = quartet_classes(args.lmax)
    rcls = [c for c in classes if is_register_class(c)]
    # register kernels
    parts = ["// Generated by gen_hgp.py -- do not edit.", "#pragma once", ""]
    for c in rcls:
        code, _, _ = emit_rkernel(c)
        parts.append(code)
        parts.append("")
    # dispatch table (class id -> register kernel or null)
    parts.append("#ifndef MI_R_BODIES_ONLY")
    parts.append("static const void* const kRegisterKernels[] = {")
    for c in classes:
        parts.append(f"  {'(const void*)hgp_r_' + cname(c) if c in rcls else 'nullptr'},  // {cname(c)}")
    parts.append("};")
    parts.append("#endif")
    with open(os.path.join(args.out, "hgp_rkernels.inc"), "w") as f:
        f.write("\n".join(parts) + "\n")
    # fused kernels: one launch over several register classes; segment s of
    # the launch takes blocks [bstart[s], bstart[s+1]) and runs its class body
    # with its own tasks and block -> task table (heaviest segments first).
    # Classes are fused within groups of similar register footprint (a fused
    # kernel takes the VGPRs of its largest body, so light classes keep their
    # occupancy), one group per translation unit (compiled in parallel).
    fz = ["// Generated by gen_hgp.py -- do not edit.", "#pragma once", "",
          f"#define HGP_NRGROUP {NRGROUP}"]
    for gi in range(NRGROUP):
        lds = any(boys_in_lds(c) for c in classes if c in rcls and register_group(c) == gi)
        fz += [f"#if defined(MI_FUSED_GROUP) && MI_FUSED_GROUP == {gi}",
               f"__global__ void __launch_bounds__(256) hgp_rf{gi}_kernel(const EriArgs args, const FusedSegs F) {{",
               "  __shared__ double sboys[BOYS_NROW * 8];" if lds else "  double* sboys = nullptr;",
               "  const int s = fused_segment(F, (int)blockIdx.x);",
               "  EriArgs a = args;",
               "  a.tasks = args.tasks + F.task_off[s];",
               "  a.ntasks = F.ntasks[s];",
               "  a.block_task = args.block_task + F.bt_off[s];",
               "  const int bid = (int)blockIdx.x - F.bstart[s];",
               "  switch (F.cls[s]) {"]
        for k, c in enumerate(classes):
            if c in rcls and register_group(c) == gi:
                fz.append(f"    case {k}: hgp_r_body_{cname(c)}(a, sboys, bid); break;")
        fz += ["    default: break;", "  }", "}", f"static const void* const kFusedRegisterKernel = (const void*)hgp_rf{gi}_kernel;",
               "#endif"]
    with open(os.path.join(args.out, "hgp_rfused.inc"), "w") as f:
        f.write("\n".join(fz) + "\n")
    # LDS-kernel tables: one contiguous blob per class, 16-byte aligned, laid
    # out as [VRR items | HRR items | out words | acc sources], so a block
    # stages it with wide loads.  Only classes whose slots fit 16 bits and whose
    # contracted block fits the per-lane accumulators (nacc <= 1024) get one;
    # boff = -1 otherwise.  Global-kernel tables (pack32) exist for every class.
    descs, BLOB, GBLOB = [], [], []
    for c in classes:
        d, vi, am, hi, ou, rx = w_tables(c)
        if d["nvrr"] < 65536 and d["nhrr"] < 65536 and d["nacc"] <= 1024:
            words = pack16(vi, hi, ou, am)
            d["boff"], d["bwords"] = len(BLOB) // 4, len(words)
            BLOB += words
        else:
            d["boff"], d["bwords"] = -1, 0
        words, offs = pack32(vi, hi, ou, am, rx)
        d["gv"], d["gh"], d["go"], d["ga"], d["gr"] = (len(GBLOB) + o for o in offs)
        GBLOB += words
        descs.append(d)
    maxlev = max(len(d["vlev"]) for d in descs)
    maxhlev = max(len(d["hlev"]) for d in descs)
    t = ["// Generated by gen_hgp.py -- do not edit.", "#pragma once", ""]
    t.append(f"#define HGP_LMAX_SHELL {args.lmax}")
    t.append(f"#define HGP_NCLASS {len(classes)}")
    t.append(f"#define HGP_MAX_VLEV {maxlev}")
    t.append(f"#define HGP_MAX_HLEV {maxhlev}")
    t.append(f"#define HGP_NRGROUP_T {NRGROUP}")
    rows = []
    for c, d in zip(classes, descs):
        vl = list(d["vlev"]) + [d["vlev"][-1]] * (maxlev - len(d["vlev"]))
        hl = list(d["hlev"]) + [d["hlev"][-1]] * (maxhlev - len(d["hlev"]))
        rows.append(f"  {{{d['la']}, {d['lb']}, {d['lc']}, {d['ld']}, {d['mmax']}, {d['nvrr']}, {d['nhrr']}, "
                 f"{d['vt0']}, {len(d['vlev']) - 1}, {len(d['hlev']) - 1}, {d['nacc']}, {d['nout']}, "
                 f"{d['boff']}, {d['bwords']}, {d['gv']}, {d['gh']}, {d['go']}, {d['ga']}, "
                 f"{{{', '.join(map(str, vl))}}}, {{{', '.join(map(str, hl))}}}, "
                 f"{1 if c in rcls else 0}, {d['gr']}}},  // {cname(c)}")
    # host copy (launch set-up) and device copy (fused wave launches read the
    # descriptor of each segment's class)
    t.append("static const WClassDesc kWClassDesc[HGP_NCLASS] = {")
    t += rows
    t.append("};")
    t.append("__constant__ WClassDesc kWClassDescDev[HGP_NCLASS] = {")
    t += rows
    t.append("};")
    t.append("// wavefronts per block of each class's register kernel: four where the")
    t.append("// block stages the Boys grid in LDS, else one (each wavefront on its own)")
    t.append("static const int kRegisterWavesPerBlock[HGP_NCLASS] = {"
             + ", ".join(str((4 if boys_in_lds(c) else 1) if c in rcls else 0) for c in classes) + "};")
    t.append("static const int kFusedWavesPerBlock[HGP_NRGROUP_T] = {"
             + ", ".join(str(4 if any(boys_in_lds(c) for c in rcls if register_group(c) == gi) else 1)
                         for gi in range(NRGROUP)) + "};")
    t.append("// fusion group of each class's register kernel (-1: none)")
    t.append("static const int kRegisterGroup[HGP_NCLASS] = {"
             + ", ".join(str(register_group(c) if c in rcls else -1) for c in classes) + "};")
    t.append("static const uint32_t kWBlob[] __attribute__((aligned(16))) = {")
    for k in range(0, len(BLOB), 12):
        t.append("  " + ", ".join(f"{x}u" for x in BLOB[k:k + 12]) + ",")
    if not BLOB:
        t.append("  0u,")
    t.append("};")
    t.append(f"static const int kWBlobWords = {len(BLOB)};")
    t.append("static const uint32_t kWGBlob[] __attribute__((aligned(16))) = {")
    for k in range(0, len(GBLOB), 16):
        t.append("  " + ",".join(f"{x}u" for x in GBLOB[k:k + 16]) + ",")
    t.append("};")
    t.append(f"static const int kWGBlobWords = {len(GBLOB)};")
    with open(os.path.join(args.out, "hgp_wtables.inc"), "w") as f:
        f.write("\n".join(t) + "\n")
    print(f"gen_hgp: {len(classes)} classes, {len(rcls)} register kernels: "
          + " ".join(cname(c) for c in rcls))


if __name__ == "__main__":
    main()
